// tn_kernels.hpp -- the relaxation's gfx950 kernels.
//
// Rebuild (src/v2m_particles.cpp:771-820): k_keys (cell key + count per cell),
// k_scan_* (exclusive scan: cell starts), k_radix_* (stable LSD radix sort of the
// nodes by cell key), k_permute (the nodes renumbered in cell order), k_truss
// (K-nearest lists, one wave per 64 nodes).
// Iteration (src/v2m_particles.cpp:828-947): k_force, k_move_easy + k_move,
// k_stats (histogram, max, Verlet count, refresh flags, FIRE power partial sums),
// k_refresh_list + k_truss_node (own-list refresh of the nodes past a full skin),
// k_control (one block: FIRE's update, p99, the stop and rebuild decisions).
//
// Node slots: at every rebuild the nodes are renumbered in cell order (slot t holds
// the node at old slot sorted[t]), so a wave's nodes, their list entries and their
// neighbours lie close in memory. id[t] is the node's index in the caller's arrays:
// it orders equal distances in the K-lists, as the reference's node index does, and
// places the results; slot_of is its inverse. The lists hold slots.
#pragma once

#include "tn_device.hpp"

namespace tnk {

using namespace tnd;

constexpr int kWave = 64;
constexpr int kBlock = 256;

// The device control block: per-iteration accumulators, FIRE's state and the
// loop decisions. The host reads its head after each iteration.
struct Ctl {
    // decisions and stats (read back by the host)
    int32_t it;            // iterations done
    int32_t done;          // converged: the loop stops
    int32_t rebuild;       // a rebuild is due before the next iteration
    int32_t rebuilds;
    float last_move, last_p99;
    int32_t fire_resets;
    float fire_dt;
    int32_t zero_v;        // FIRE reset: the next move reads every velocity as zero
    int32_t power_fallbacks;   // iterations whose power sign needed the serial sum
    int32_t nrefresh;      // nodes queued for an own-list refresh this iteration
    int32_t nhalf;         // nodes past skin/2 (this iteration)
    int32_t nmove;         // interface nodes queued for the trapped step (this iteration)
    int32_t nmove_int;     // interior nodes queued for the trapped step (crossing or in snap range)
    int32_t nmove_rest;    // other nodes queued for the general trapped step (this iteration)
    int32_t nmove_def;     // nodes a lean pass handed to the general step (this iteration)
    // FIRE (FireCtl, src/v2m_particles.h:47-76)
    float dt0, dt, dtmax, alpha;
    int32_t npos;
    // accumulators, cleared by k_control
    uint32_t hist[32];
    uint32_t mmax_bits;
    int32_t order_cnt[4];  // node counts by type rank (interface, junction, interior, corner)
    double power;
};

// ---- small utilities ------------------------------------------------------------

__global__ void k_fill_i32(int32_t* __restrict__ a, int64_t m, int32_t v) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) {
        a[i] = v;
    }
}

__device__ __forceinline__ uint64_t lanemask_lt() {
    const int lane = threadIdx.x & (kWave - 1);
    return lane ? (~0ull >> (kWave - lane)) : 0ull;
}

// exclusive scan of one int per thread over a 256-thread block; *total = block sum
__device__ __forceinline__ int block_excl_scan(int v, int* lds4, int* total) {
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x >> 6;
    int x = v;
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        const int y = __shfl_up(x, d, kWave);
        if (lane >= d) {
            x += y;
        }
    }
    if (lane == kWave - 1) {
        lds4[w] = x;
    }
    __syncthreads();
    int base = 0, tot = 0;
#pragma unroll
    for (int k = 0; k < kBlock / kWave; ++k) {
        const int s = lds4[k];
        base += k < w ? s : 0;
        tot += s;
    }
    __syncthreads();
    *total = tot;
    return base + x - v;
}

// ---- exclusive scan (int32), tiles of 4096 ----------------------------------------

constexpr int kScanItems = 16;
constexpr int kScanTile = kBlock * kScanItems;

__global__ __launch_bounds__(kBlock) void k_scan_reduce(const int32_t* __restrict__ in, int64_t m,
                                                        int32_t* __restrict__ bsum) {
    __shared__ int lds4[4];
    const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
    int s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        s += base + k < m ? in[base + k] : 0;
    }
    int tot;
    block_excl_scan(s, lds4, &tot);
    if (threadIdx.x == 0) {
        bsum[blockIdx.x] = tot;
    }
}

// one block: exclusive scan of the tile sums in place, chunk by chunk
__global__ __launch_bounds__(kBlock) void k_scan_top(int32_t* __restrict__ bsum, int64_t nb) {
    __shared__ int lds4[4];
    int carry = 0;
    for (int64_t c0 = 0; c0 < nb; c0 += kScanTile) {
        const int64_t base = c0 + (int64_t)threadIdx.x * kScanItems;
        int v[kScanItems];
        int s = 0;
#pragma unroll
        for (int k = 0; k < kScanItems; ++k) {
            v[k] = base + k < nb ? bsum[base + k] : 0;
            s += v[k];
        }
        int tot;
        int run = carry + block_excl_scan(s, lds4, &tot);
#pragma unroll
        for (int k = 0; k < kScanItems; ++k) {
            if (base + k < nb) {
                bsum[base + k] = run;
            }
            run += v[k];
        }
        carry += tot;
    }
}

__global__ __launch_bounds__(kBlock) void k_scan_down(const int32_t* in, int32_t* out, int64_t m,
                                                      const int32_t* __restrict__ bsum) {
    __shared__ int lds4[4];
    const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
    int v[kScanItems];
    int s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        v[k] = base + k < m ? in[base + k] : 0;
        s += v[k];
    }
    int tot;
    int run = bsum[blockIdx.x] + block_excl_scan(s, lds4, &tot);
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        if (base + k < m) {
            out[base + k] = run;
        }
        run += v[k];
    }
}

// ---- keys (v2m_h_at + v2m_level_of + v2m_bin_key; count per cell) ---------------

// X.w is v2m_h_at at the node (set with every move), the size the rebuild takes.
// The nodes sit in cell order from the previous rebuild, so neighbouring lanes
// mostly share a key: each run of equal keys in a wave adds its length at once.
__global__ __launch_bounds__(kBlock) void k_keys(Hash H, const float4* __restrict__ X, int n, float tskin,
                                                 int32_t* __restrict__ key, int32_t* __restrict__ cnt) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    const int lane = threadIdx.x & (kWave - 1);
    const bool valid = i < n;
    int k = -1;
    if (valid) {
        const float4 x = X[i];
        k = cell_key(H, level_for(H, tskin * x.w), x.x, x.y, x.z);
        key[i] = k;
    }
    const int prev = __shfl_up(k, 1, kWave);
    const bool head = valid && (lane == 0 || prev != k);
    const uint64_t heads = __ballot(head);
    const int nvalid = __popcll(__ballot(valid));   // the valid lanes are a prefix
    if (head) {
        const uint64_t later = lane == kWave - 1 ? 0ull : heads & (~0ull << (lane + 1));
        const int end = later ? __ffsll((long long)later) - 1 : nvalid;
        atomicAdd(&cnt[k], end - lane);
    }
}

// ---- the K-nearest search's query order ------------------------------------------
//
// k_truss serves the nodes a wave at a time. In slot order (cells in x-fastest row
// order) a wave's 64 nodes lie along a row of cells, and the union of their scan
// boxes is long; taken brick by brick instead (4 x 4 x 4 cells of a level, cells
// z, y, x within a brick, bricks in x-fastest order) they are compact, and the union,
// which a wave loads and tests as a whole, is about a quarter smaller. The order is a
// counting sort of the slots by brick position: a cell's slots stay contiguous.

constexpr int kBrickEdge = 4;

struct BrickMap {
    int nb[TN_MAXLEV][3];   // bricks per axis at each level
    int off[TN_MAXLEV];     // each level's first position (64 a brick)
    int size;               // positions in all
};

// the brick-order position of cell key k
__device__ __forceinline__ int brick_pos(const Hash& H, const BrickMap& B, int k) {
    int L = 0;
    while (L + 1 < H.nlev && k >= H.off[L + 1]) {
        ++L;
    }
    const int local = k - H.off[L], nx = H.dim[L][0], ny = H.dim[L][1];
    const int cx = local % nx, cy = (local / nx) % ny, cz = local / (nx * ny);
    const int brick = (cx / kBrickEdge) + B.nb[L][0] * ((cy / kBrickEdge) + B.nb[L][1] * (cz / kBrickEdge));
    const int within = (cx % kBrickEdge) + kBrickEdge * ((cy % kBrickEdge) + kBrickEdge * (cz % kBrickEdge));
    return B.off[L] + brick * (kBrickEdge * kBrickEdge * kBrickEdge) + within;
}

// the node count of each cell at its brick-order position (positions of no cell stay 0)
__global__ __launch_bounds__(kBlock) void k_brick_count(Hash H, BrickMap B, const int32_t* __restrict__ cstart,
                                                        int32_t* __restrict__ cnt) {
    const int k = blockIdx.x * kBlock + threadIdx.x;
    if (k < H.nkeys) {
        cnt[brick_pos(H, B, k)] = cstart[k + 1] - cstart[k];
    }
}

// the slots in brick order: slot s goes to its cell's start in that order plus its rank
// in the cell (start: the exclusive scan of k_brick_count's counts)
__global__ __launch_bounds__(kBlock) void k_brick_order(Hash H, BrickMap B, const int32_t* __restrict__ key, int n,
                                                        const int32_t* __restrict__ cstart,
                                                        const int32_t* __restrict__ start,
                                                        int32_t* __restrict__ order) {
    const int s = blockIdx.x * kBlock + threadIdx.x;
    if (s < n) {
        const int k = key[s];
        order[start[brick_pos(H, B, k)] + (s - cstart[k])] = s;
    }
}

// ---- stable LSD radix sort of (key, id) by key, 8-bit digits --------------------------

constexpr int kRadixItems = 8;
constexpr int kRadixTile = kBlock * kRadixItems;

__global__ __launch_bounds__(kBlock) void k_radix_hist(const int32_t* __restrict__ keys, int n, int shift, int nblk,
                                                       int32_t* __restrict__ bhist) {
    __shared__ int lh[256];
    lh[threadIdx.x] = 0;
    __syncthreads();
    const int base = blockIdx.x * kRadixTile;
#pragma unroll
    for (int it = 0; it < kRadixItems; ++it) {
        const int e = base + it * kBlock + threadIdx.x;
        if (e < n) {
            atomicAdd(&lh[(keys[e] >> shift) & 255], 1);
        }
    }
    __syncthreads();
    bhist[(size_t)threadIdx.x * nblk + blockIdx.x] = lh[threadIdx.x];
}

// Scatter one tile: element order within the tile is (round, thread), and each
// element's rank among equal digits counts the earlier ones -- a stable pass.
// vals_in == nullptr: the values are the element indices (the node ids).
__global__ __launch_bounds__(kBlock) void k_radix_scatter(const int32_t* __restrict__ keys_in,
                                                          const int32_t* __restrict__ vals_in, int n, int shift,
                                                          int nblk, const int32_t* __restrict__ offs,
                                                          int32_t* __restrict__ keys_out,
                                                          int32_t* __restrict__ vals_out) {
    __shared__ int goff[256];
    __shared__ int run[256];
    __shared__ int wcnt[kBlock / kWave][256];
    const int tid = threadIdx.x, w = tid >> 6;
    goff[tid] = offs[(size_t)tid * nblk + blockIdx.x];
    run[tid] = 0;
    const int base = blockIdx.x * kRadixTile;
    const uint64_t below = lanemask_lt();
    for (int it = 0; it < kRadixItems; ++it) {
        const int e = base + it * kBlock + tid;
        const bool valid = e < n;
        const int key = valid ? keys_in[e] : 0;
        const int val = valid ? (vals_in ? vals_in[e] : e) : 0;
        const int d = valid ? ((key >> shift) & 255) : 256;
        // lanes of this wave holding the same digit (9 bits: 256 = no element)
        uint64_t peers = ~0ull;
#pragma unroll
        for (int b = 0; b < 9; ++b) {
            const uint64_t bal = __ballot((d >> b) & 1);
            peers &= ((d >> b) & 1) ? bal : ~bal;
        }
        const int rank_w = __popcll(peers & below);
#pragma unroll
        for (int k = 0; k < kBlock / kWave; ++k) {
            wcnt[k][tid] = 0;
        }
        __syncthreads();
        if (valid && rank_w == 0) {
            wcnt[w][d] = __popcll(peers);
        }
        __syncthreads();
        {
            int acc = run[tid];
#pragma unroll
            for (int k = 0; k < kBlock / kWave; ++k) {
                const int c = wcnt[k][tid];
                wcnt[k][tid] = acc;
                acc += c;
            }
            run[tid] = acc;
        }
        __syncthreads();
        if (valid) {
            const int pos = goff[d] + wcnt[w][d] + rank_w;
            keys_out[pos] = key;
            vals_out[pos] = val;
        }
        __syncthreads();
    }
}

// ---- renumbering in cell order ---------------------------------------------------------

// Slot t takes the node of old slot sorted[t]; Ps and P0 (the positions as of this
// rebuild) and Ms (id, typ | lab << 8) are written in the new order.
struct NodeArrays {
    float4* X;
    float4* V;
    uint8_t* typ;
    uint32_t* part;
    uint16_t* lab;
    int32_t* id;
    uint8_t* sticky;   // handed back to the general move pass before (may be null)
};

__global__ __launch_bounds__(kBlock) void k_permute(const int32_t* __restrict__ sorted, int n, NodeArrays in,
                                                    NodeArrays out, float4* __restrict__ Ps,
                                                    float4* __restrict__ P0, int2* __restrict__ Ms,
                                                    int32_t* __restrict__ slot_of) {
    const int t = blockIdx.x * kBlock + threadIdx.x;
    if (t < n) {
        const int s = sorted[t];
        const float4 x = in.X[s];
        out.X[t] = x;
        Ps[t] = x;
        P0[t] = x;
        out.V[t] = in.V[s];
        const int ty = in.typ[s];
        out.typ[t] = (uint8_t)ty;
        out.part[t] = in.part[s];
        const int l = in.lab[s];
        out.lab[t] = (uint16_t)l;
        const int o = in.id[s];
        out.id[t] = o;
        if (in.sticky) {
            out.sticky[t] = in.sticky[s];
        }
        Ms[t] = make_int2(o, ty | (l << 8));
        slot_of[o] = t;
    }
}

// ---- K-nearest lists, one wave per 64 nodes in cell order ----------------------------
//
// v2m_neighbors (src/opencl/v2m_particle_body.cl:73-151) for 64 query nodes at once.
// The queries are consecutive in cell order, so their scan boxes overlap. Per group
// of lanes (one cell row, a bounded span along it) the wave lists the rows of the
// union of its lanes' boxes over their levels (rows no lane reaches are skipped),
// reads all their cell starts at once, and tests the concatenation of the rows'
// candidates in full chunks of 64 staged in LDS: a distance prefilter on the
// matrix cores, then for every possible hit of a lane that scans the candidate's
// row the reference's exact test and its other checks (not the node itself, no
// interior bar across labels). Valid candidates are appended as (distance, id) to
// the lane's LDS buffer, which is cut back to its 24 best when it fills; at the end
// the first 24 in (distance, id) order are kept. The list equals the reference's:
// its K smallest valid candidates in (distance, id) order, and every valid
// candidate lies inside the lane's own box, which the union contains. (That holds
// when every node's reach fits its level's cell, H.prune; otherwise a lane also
// checks that the candidate's cell is in its own box, the reference's scan region,
// before accepting it.)

constexpr int kTrussWave = 64;   // one wave per block
constexpr int kCollect = 32;   // per-lane candidate buffer (entries): settled by a 32-key network
constexpr int kMinRoom = 4;    // a lane with less room than this (and more possible hits) settles first

// ---- wave-wide reductions and scans on DPP (64 lanes, all active) ----------------
//
// Each step combines a lane with a lane of its row (16 lanes) through the data-
// parallel-primitives operand modifier, then rows with row broadcasts (lane 15 of a
// row into the next, lane 31 into the upper two), so the reduction never touches LDS.
// dpp_ctrl: quad_perm [1,0,3,2] 0xB1, [2,3,0,1] 0x4E; row_shr:k 0x110 + k;
// row_mirror 0x140, row_half_mirror 0x141; row_bcast:15 0x142, row_bcast:31 0x143.

template <int kCtrl, int kRowMask = 0xF>
__device__ __forceinline__ int dpp(int old, int v) {
    return __builtin_amdgcn_update_dpp(old, v, kCtrl, kRowMask, 0xF, false);
}

// the reduction of v over the wave under op (commutative, idempotent), as a uniform value
template <typename Op>
__device__ __forceinline__ int wave_reduce(int v, Op op) {
    v = op(v, dpp<0xB1>(v, v));
    v = op(v, dpp<0x4E>(v, v));
    v = op(v, dpp<0x141>(v, v));
    v = op(v, dpp<0x140>(v, v));
    v = op(v, dpp<0x142, 0xA>(v, v));
    v = op(v, dpp<0x143, 0xC>(v, v));
    return __builtin_amdgcn_readlane(v, kWave - 1);
}

__device__ __forceinline__ int wave_min(int v) {
    return wave_reduce(v, [](int a, int b) { return min(a, b); });
}

__device__ __forceinline__ int wave_max(int v) {
    return wave_reduce(v, [](int a, int b) { return max(a, b); });
}

// the largest of non-negative floats (their bits order as unsigned integers do)
__device__ __forceinline__ float wave_max_nonneg(float v) {
    return __uint_as_float((uint32_t)wave_reduce((int)__float_as_uint(v), [](int a, int b) {
        return (int)max((uint32_t)a, (uint32_t)b);
    }));
}

// exclusive prefix sum over the wave; *total: the sum of all lanes
__device__ __forceinline__ int wave_excl_scan(int v, int lane, int* total) {
    (void)lane;
    int x = v;
    x += dpp<0x111>(0, x);   // row_shr: a lane with no source adds 0
    x += dpp<0x112>(0, x);
    x += dpp<0x114>(0, x);
    x += dpp<0x118>(0, x);
    x += dpp<0x142, 0xA>(0, x);
    x += dpp<0x143, 0xC>(0, x);
    *total = __builtin_amdgcn_readlane(x, kWave - 1);
    return x - v;
}

// (distance, id) order of the K-list (src/opencl/v2m_particle_body.cl:128-141)
__device__ __forceinline__ bool before(float da, int ja, float db, int jb) {
    return da < db || (da == db && ja < jb);
}

// A list entry as one 64-bit key: the distance's bits above the id. Distances are
// >= +0, whose bits order as the floats do, so the keys order as (distance, id), the
// K-list's order (src/opencl/v2m_particle_body.cl:128-141).
__device__ __forceinline__ uint64_t list_key(float d, int id) {
    return ((uint64_t)__float_as_uint(d) << 32) | (uint32_t)id;
}

// Sort 32 keys in registers: a bitonic network, every index known at compile time.
__device__ __forceinline__ void bitonic32(uint64_t (&k)[32]) {
#pragma unroll
    for (int size = 2; size <= 32; size <<= 1) {
#pragma unroll
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
#pragma unroll
            for (int i = 0; i < 32; ++i) {
                const int j = i ^ stride;
                if (j > i) {
                    const uint64_t x = k[i], y = k[j];
                    const bool swap = (i & size) == 0 ? x > y : x < y;
                    k[i] = swap ? y : x;
                    k[j] = swap ? x : y;
                }
            }
        }
    }
}

// ---- settling a few lanes' buffers with the whole wave -----------------------------------
//
// When only some lanes' buffers are full, the wave sorts them two at a time, one per
// half-wave: lane i of a half holds entry i of its buffer (the key, with the slot it
// came from) and a bitonic network runs across the 32 lanes, each compare-exchange
// with partner lane i ^ j taken through DPP (quad_perm for j = 1, 2; row_ror:4 and
// row_ror:12 for j = 4; row_ror:8 for j = 8) or ds_swizzle's bit mode (j = 16, xor
// within 32 lanes). 15 steps of a few instructions against a 32-key register network
// per settling lane.

template <int kCtrl>
__device__ __forceinline__ uint64_t dpp64(uint64_t v) {
    const int lo = (int)(uint32_t)v, hi = (int)(uint32_t)(v >> 32);
    return ((uint64_t)(uint32_t)dpp<kCtrl>(hi, hi) << 32) | (uint32_t)dpp<kCtrl>(lo, lo);
}

// the value of lane (i ^ J) of this lane's half-wave, i = lane % 32
template <int J>
__device__ __forceinline__ uint64_t xor_partner(uint64_t v, int i) {
    if constexpr (J == 1) {
        return dpp64<0xB1>(v);
    } else if constexpr (J == 2) {
        return dpp64<0x4E>(v);
    } else if constexpr (J == 4) {
        const uint64_t dn = dpp64<0x124>(v), up = dpp64<0x12C>(v);
        return (i & 4) ? dn : up;
    } else if constexpr (J == 8) {
        return dpp64<0x128>(v);
    } else {
        static_assert(J == 16, "xor partners within 32 lanes");
        const int lo = __builtin_amdgcn_ds_swizzle((int)(uint32_t)v, 0x401F);
        const int hi = __builtin_amdgcn_ds_swizzle((int)(uint32_t)(v >> 32), 0x401F);
        return ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo;
    }
}

// the 32-bit value of lane (i ^ J) of this lane's half-wave
template <int J>
__device__ __forceinline__ int xor_partner32(int v, int i) {
    if constexpr (J == 1) {
        return dpp<0xB1>(v, v);
    } else if constexpr (J == 2) {
        return dpp<0x4E>(v, v);
    } else if constexpr (J == 4) {
        const int dn = dpp<0x124>(v, v), up = dpp<0x12C>(v, v);
        return (i & 4) ? dn : up;
    } else if constexpr (J == 8) {
        return dpp<0x128>(v, v);
    } else {
        return __builtin_amdgcn_ds_swizzle(v, 0x401F);
    }
}

// one step of the network on (key, payload); keys are distinct or both padding
template <int K, int J>
__device__ __forceinline__ void half_bitonic_step_kv(uint64_t& v, int& pl, int i) {
    const uint64_t o = xor_partner<J>(v, i);
    const int op = xor_partner32<J>(pl, i);
    const bool keep_min = ((i & J) == 0) == (K >= 32 || (i & K) == 0);
    const bool take = keep_min == (o < v);
    v = take ? o : v;
    pl = take ? op : pl;
}

__device__ __forceinline__ void half_wave_sort32_kv(uint64_t& v, int& pl, int lane) {
    const int i = lane & 31;
    half_bitonic_step_kv<2, 1>(v, pl, i);
    half_bitonic_step_kv<4, 2>(v, pl, i);
    half_bitonic_step_kv<4, 1>(v, pl, i);
    half_bitonic_step_kv<8, 4>(v, pl, i);
    half_bitonic_step_kv<8, 2>(v, pl, i);
    half_bitonic_step_kv<8, 1>(v, pl, i);
    half_bitonic_step_kv<16, 8>(v, pl, i);
    half_bitonic_step_kv<16, 4>(v, pl, i);
    half_bitonic_step_kv<16, 2>(v, pl, i);
    half_bitonic_step_kv<16, 1>(v, pl, i);
    half_bitonic_step_kv<32, 16>(v, pl, i);
    half_bitonic_step_kv<32, 8>(v, pl, i);
    half_bitonic_step_kv<32, 4>(v, pl, i);
    half_bitonic_step_kv<32, 2>(v, pl, i);
    half_bitonic_step_kv<32, 1>(v, pl, i);
}

__device__ __forceinline__ uint64_t readlane64(uint64_t v, int l) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, l);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), l);
    return ((uint64_t)hi << 32) | lo;
}

// A slot buffer's entry e of lane `who` as its key, recomputed from the candidate's
// position as the test formed it: (distance from the query at (x, y, z), id).
__device__ __forceinline__ uint64_t slot_key(const float4* __restrict__ Ps, const int2* __restrict__ Ms, int slot,
                                             float x, float y, float z) {
    const float4 c = Ps[slot];
    const float ex = c.x - x, ey = c.y - y, ez = c.z - z;
    const float d2 = ex * ex + ey * ey + ez * ez;
    return list_key(sqrtf(d2), Ms[slot].x);
}

__device__ __forceinline__ float readlane_f(float v, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

// Cut the slot buffers of the lanes in sm (wave-uniform) back to their 24 best (in
// order); their cnt becomes 24 and kth the 24th key. All lanes take part: the keys are
// recomputed from the slots, relative to the owner's query (px, py, pz).
__device__ __forceinline__ void settle_slot_lanes(uint32_t* cb, int lane, uint64_t sm, int& cnt, uint64_t& kth,
                                                  const float4* __restrict__ Ps, const int2* __restrict__ Ms,
                                                  float px, float py, float pz) {
    while (sm) {
        const int a = __builtin_ctzll(sm);
        sm &= sm - 1;
        const int b = sm ? __builtin_ctzll(sm) : a;
        sm &= sm ? sm - 1 : 0ull;
        const int ca = __builtin_amdgcn_readlane(cnt, a), cb_n = __builtin_amdgcn_readlane(cnt, b);
        const float xa = readlane_f(px, a), ya = readlane_f(py, a), za = readlane_f(pz, a);
        const float xb = readlane_f(px, b), yb = readlane_f(py, b), zb = readlane_f(pz, b);
        const bool lo = lane < 32;
        const int who = lo ? a : b, cw = lo ? ca : cb_n;
        const int i = lane & 31;
        int sl = 0;
        uint64_t v = ~0ull;
        if (i < cw) {
            sl = (int)cb[i * kTrussWave + who];
            v = slot_key(Ps, Ms, sl, lo ? xa : xb, lo ? ya : yb, lo ? za : zb);
        }
        half_wave_sort32_kv(v, sl, lane);
        if (i < kK) {
            cb[i * kTrussWave + who] = (uint32_t)sl;
        }
        const uint64_t ka = readlane64(v, kK - 1), kb = readlane64(v, 32 + kK - 1);
        if (lane == a) {
            cnt = kK;
            kth = ka;
        }
        if (lane == b) {
            cnt = kK;
            kth = kb;
        }
    }
}

// ---- the distance prefilter on the matrix cores ------------------------------------------
//
// For query p (this lane's node) and candidate c, with both taken relative to a
// wave-uniform origin and k = (t + skin) / 2, the exact test is d^2 < (k (h_p + h_c))^2.
// Its difference d^2 - k^2 (h_p + h_c)^2 is a dot product of two 6-vectors,
//   b_c = (-2c, -2 k^2 h_c, 1, |c|^2 - k^2 h_c^2),  a_p = (p, h_p, |p|^2 - k^2 h_p^2, 1),
// so the 64 x 64 differences of a chunk of candidates and a wave's queries come from
// four 32 x 32 tiles of v_mfma_f32_32x32x2f32 (three K-steps each), candidates along
// the rows and queries along the columns. A candidate whose difference is below a
// rounding margin is a possible hit; every possible hit is then tested exactly as
// before, so the prefilter only has to keep every true hit.
// (v_mfma_f32_32x32x2f32: lane l supplies A[l % 32][l / 32] and B[l / 32][l % 32];
// its output register e holds D[8 (e / 4) + 4 (l / 32) + e % 4][l % 32]. So in the
// tile of its own query's columns a lane holds half of the tile's candidate rows
// for its query, and lane l ^ 32 holds the other half.)

typedef float f32x16 __attribute__((ext_vector_type(16)));


// this lane's B operands: queries 32 c + l % 32 (c = the column tile), three K-steps
struct QueryOps {
    float b[2][3];
};

__device__ __forceinline__ QueryOps query_ops(float qx, float qy, float qz, float h, float Aq, int lane) {
    const bool up = lane >= 32;
    QueryOps o;
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        const int src = 32 * c + (lane & 31);
        const float x = __shfl(qx, src, kWave), y = __shfl(qy, src, kWave), z = __shfl(qz, src, kWave);
        const float hh = __shfl(h, src, kWave), A = __shfl(Aq, src, kWave);
        // K-index 2 t + l / 32: (x, y), (z, h), (|p|^2 - k^2 h^2, 1)
        o.b[c][0] = up ? y : x;
        o.b[c][1] = up ? hh : z;
        o.b[c][2] = up ? 1.0f : A;
    }
    return o;
}

// The tests of one output tile, dense: bit e for register e (its difference below the
// margin). Plain compares, so that the compiler keeps the wait states the matrix
// cores' results need before a vector instruction may read them.
__device__ __forceinline__ uint32_t tile_bits(const f32x16& acc, float margin) {
    uint32_t m = 0;
#pragma unroll
    for (int e = 0; e < 16; ++e) {
        m |= (acc[e] < margin ? 1u : 0u) << e;
    }
    return m;
}

// the 16 bits of a dense tile mask spread to nibbles at bits 8 k .. 8 k + 3
__device__ __forceinline__ uint32_t spread_nibbles(uint32_t d) {
    uint32_t x = d & 0xFFFFu;
    x = (x | (x << 8)) & 0x00FF00FFu;
    return (x | (x << 4)) & 0x0F0F0F0Fu;
}

// the 64-bit mask of this lane's query over the chunk's candidates (cpos) whose
// difference, relative to the origin o, is below `margin`
__device__ __forceinline__ uint64_t chunk_hits(const QueryOps& qo, const float4* cpos, float ox, float oy, float oz,
                                               float k2, float margin, int lane) {
    const bool up = lane >= 32;
    const int h = lane >> 5;
    uint32_t out[2];
#pragma unroll
    for (int r = 0; r < 2; ++r) {   // candidates 32 r .. 32 r + 31 (the row tile)
        const float4 v = cpos[32 * r + (lane & 31)];
        float a0, a1, a2;
        if (v.x < __builtin_huge_valf()) {
            const float cx = v.x - ox, cy = v.y - oy, cz = v.z - oz;
            const float C = cx * cx + cy * cy + cz * cz - k2 * v.w * v.w;
            // K-index 2 t + l / 32: (-2 cx, -2 cy), (-2 cz, -2 k^2 h), (1, |c|^2 - k^2 h^2)
            a0 = up ? -2.0f * cy : -2.0f * cx;
            a1 = up ? -2.0f * k2 * v.w : -2.0f * cz;
            a2 = up ? C : 1.0f;
        } else {   // past the row's end: a difference far above any margin
            a0 = 0.0f;
            a1 = 0.0f;
            a2 = up ? 1e30f : 0.0f;
        }
        uint32_t bits[2];
#pragma unroll
        for (int c = 0; c < 2; ++c) {   // queries 32 c .. 32 c + 31 (the column tile)
            f32x16 acc = {};
            acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a0, qo.b[c][0], acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a1, qo.b[c][1], acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a2, qo.b[c][2], acc, 0, 0, 0);
            bits[c] = tile_bits(acc, margin);
        }
        // my query's column tile is h: register e of lane l holds candidate
        // 8 (e / 4) + 4 (l / 32) + e % 4 of the row tile; lane l ^ 32 has the other half
        const uint32_t mine = h ? bits[1] : bits[0], give = h ? bits[0] : bits[1];
        const uint32_t got = (uint32_t)__shfl_xor((int)give, 32, kWave);
        out[r] = (spread_nibbles(mine) << (4 * h)) | (spread_nibbles(got) << (4 * (1 - h)));
    }
    return ((uint64_t)out[1] << 32) | out[0];
}

// The rows of one group, collected before any of their candidates is loaded: the
// cell-start reads of all rows are then issued together, and the candidates of all
// rows (and levels) are concatenated into full chunks.
constexpr int kMaxRanges = 64;


// The lists of all nodes at a rebuild: query q is the node in slot q. The buffers hold
// the candidates' slots (4 bytes) and the keys are recomputed from them when a buffer
// is cut back and at the end: 12 KB of LDS a wave (instead of 20 KB with the keys
// themselves), so that 12 waves fit a CU.
__global__ __launch_bounds__(kTrussWave, 3) void k_truss(Hash H, TrussIn in, int n, int32_t* __restrict__ nbrT,
                                                         int32_t* __restrict__ nnb) {
    __shared__ float4 cpos[kTrussWave];   // the chunk's candidates: position, h
    __shared__ int2 cms[kTrussWave];      // their (id, typ | lab << 8)
    __shared__ int crow[kTrussWave];      // their row (index into the row table)
    // the row table of the current group: cell-start indices of the row's first cell and
    // one past its last, the lanes that scan the row, its level; after the cell starts
    // are read: the row's first slot and its position in the concatenation
    __shared__ int r_key0[kMaxRanges], r_key1[kMaxRanges], r_lev[kMaxRanges], r_sa[kMaxRanges], r_beg[kMaxRanges],
        r_len[kMaxRanges];
    // a round's (query lane, candidate) pairs; the row keys are read before any round
    static_assert(kMaxRanges >= kTrussWave, "the pairs share the row keys' array");
    int* const pairs = r_key0;
    __shared__ uint64_t r_mask[kMaxRanges];
    __shared__ uint32_t cb[kCollect * kTrussWave];   // the lanes' lists (candidate slots)
    __shared__ int cslot[kTrussWave];                // the chunk's candidates' slots
    const int lane = threadIdx.x;
    const float tskin = in.t + in.skin;
    const float half = tskin * 0.5f;   // (t + skin) * 0.5f * (hi + hj), left to right
    const float k2 = half * half;
    const float inf = __builtin_huge_valf();
    for (int qb = blockIdx.x * kTrussWave; qb < n; qb += gridDim.x * kTrussWave) {
        const bool active = qb + lane < n;
        const int q = active ? in.order[qb + lane] : 0;   // the node in slot q
        float px = 0.0f, py = 0.0f, pz = 0.0f, hi = 0.0f, Ri = 0.0f;
        int Li = 0, ti_int = 0, li = 0, myid = -1;
        if (active) {   // the node in slot q (nodes are in cell order)
            const float4 xi = in.X[q];
            px = xi.x;
            py = xi.y;
            pz = xi.z;
            hi = xi.w;
            Ri = tskin * hi;
            Li = level_for(H, Ri);
            const int2 mi = in.Ms[q];
            myid = mi.x;
            ti_int = (mi.y & 0xFF) == TN_INTERIOR;
            li = mi.y >> 8;
        }
        // The lanes are served in groups that share their level and a row of bricks and lie
        // within one brick of the group's first lane along it, so that a group's union of
        // scan boxes stays tight (a wave of queries can straddle bricks).
        int grow = -1, gx = 0;
        if (active) {
            const int cy = cell_coord(H, Li, py, H.oy, 1), cz = cell_coord(H, Li, pz, H.oz, 2);
            grow = (Li << 26) | ((cz / kBrickEdge) << 13) | (cy / kBrickEdge);
            gx = cell_coord(H, Li, px, H.ox, 0) / kBrickEdge;
        }
        bool pending = active;
        // The lane's buffer holds cnt keys; once it has been cut back to the 24 best, kth is
        // the 24th of them, and a candidate must come before it to enter.
        int cnt = 0;
        uint64_t kth = ~0ull;
        for (;;) {
            const uint64_t pm = __ballot(pending);
            if (!pm) {
                break;
            }
            const int lead = __ffsll((long long)pm) - 1;
            const int lrow = __shfl(grow, lead, kWave), lx = __shfl(gx, lead, kWave);
            const bool ingroup = pending && grow == lrow && abs(gx - lx) <= 1;
            pending = pending && !ingroup;
            // the prefilter's origin: the group's first query; its query operands
            const float gox = __shfl(px, lead, kWave), goy = __shfl(py, lead, kWave), goz = __shfl(pz, lead, kWave);
            const float qx = px - gox, qy = py - goy, qz = pz - goz;
            const float q2 = qx * qx + qy * qy + qz * qz;
            const QueryOps qops = query_ops(qx, qy, qz, hi, q2 - k2 * hi * hi, lane);
            const float P2 = wave_max_nonneg(ingroup ? q2 : 0.0f), HP = wave_max_nonneg(ingroup ? hi : 0.0f);
            int nr = 0;   // rows in the table (wave-uniform)
            // Read the cell starts of the table's rows, then test the concatenation of their
            // candidates chunk by chunk: the distance prefilter on the matrix cores, then the
            // exact test and the reference's other checks for each possible hit of a lane
            // that scans the candidate's row; the valid ones are appended as (distance, id).
            auto flush = [&]() {
                __syncthreads();   // the table's rows are written
                int sa = 0, len = 0;
                if (lane < nr) {
                    sa = in.cstart[r_key0[lane]];
                    len = in.cstart[r_key1[lane]] - sa;
                }
                int total = 0;
                const int beg = wave_excl_scan(len, lane, &total);
                if (lane < nr) {
                    r_sa[lane] = sa;
                    r_beg[lane] = beg;
                    r_len[lane] = len;
                }
                const int rows = nr;
                nr = 0;
                __syncthreads();
                // The candidates of chunk base + 64 are loaded while chunk base is tested.
                int rfirst = 0;   // the row holding the loaded chunk's first candidate
                float4 npos;
                int2 nms;
                int nrow, nslot;
                auto fetch = [&](int base) {
                    while (rfirst + 1 < rows && r_beg[rfirst + 1] <= base) {
                        ++rfirst;
                    }
                    const int p = base + lane;
                    int lo = rfirst;   // the candidate's row: the last that starts at or before p
                    if (p < total) {
                        while (lo + 1 < rows && r_beg[lo + 1] <= p) {
                            ++lo;
                        }
                        const int s = r_sa[lo] + (p - r_beg[lo]);
                        npos = in.Ps[s];
                        nms = in.Ms[s];
                        nslot = s;
                    } else {   // past the end: far away, it fails every test
                        npos = make_float4(inf, inf, inf, 0.0f);
                        nms = make_int2(-1, 0);
                        nslot = 0;
                    }
                    nrow = lo;
                };
                fetch(0);
                for (int base = 0; base < total; base += kTrussWave) {
                    __syncthreads();   // the previous chunk is consumed
                    cpos[lane] = npos;
                    cms[lane] = nms;
                    crow[lane] = nrow;
                    cslot[lane] = nslot;
                    const int lo = nrow;
                    __syncthreads();
                    if (base + kTrussWave < total) {
                        fetch(base + kTrussWave);
                    }
                    const int cn = min(kTrussWave, total - base);
                    // the chunk's own bound for the rounding margin
                    const float4 mine_c = cpos[lane];
                    const bool cval = mine_c.x < inf;
                    const float ccx = mine_c.x - gox, ccy = mine_c.y - goy, ccz = mine_c.z - goz;
                    const float C2 = wave_max_nonneg(cval ? ccx * ccx + ccy * ccy + ccz * ccz : 0.0f);
                    const float HC = wave_max_nonneg(cval ? mine_c.w : 0.0f);
                    const float hs = HP + HC;
                    const float margin = (2.0f * (P2 + C2) + 2.0f * k2 * hs * hs) * (1.0f / 65536.0f);
                    uint64_t hit = chunk_hits(qops, cpos, gox, goy, goz, k2, margin, lane);
                    // the candidates of the rows this lane scans (its level range, its own box)
                    uint64_t elig = 0;
                    {
                        const int r0 = __shfl(lo, 0, kWave), r1 = __shfl(lo, cn - 1, kWave);
                        for (int r = r0; r <= r1; ++r) {
                            const int a = max(r_beg[r] - base, 0), b = min(r_beg[r] + r_len[r] - base, cn);
                            if (b > a && ((r_mask[r] >> lane) & 1ull)) {
                                elig |= (b - a == kTrussWave ? ~0ull : ((1ull << (b - a)) - 1)) << a;
                            }
                        }
                    }
                    hit = ingroup ? hit & elig : 0ull;
                    // The possible hits as (query lane, candidate) pairs, up to 64 a round, one
                    // a lane: the exact test and the reference's other checks, and the valid
                    // ones appended to their query's buffer, packed in pair order.
                    while (__ballot(hit != 0ull)) {
                        const int pc = __popcll(hit);
                        // little room for them: keep the 24 best so far
                        const uint64_t sm = __ballot(pc > kCollect - cnt && kCollect - cnt < kMinRoom);
                        if (sm) {
                            settle_slot_lanes(cb, lane, sm, cnt, kth, in.Ps, in.Ms, px, py, pz);
                        }
                        const int want = min(pc, kCollect - cnt);
                        int npair = 0;
                        const int s0 = wave_excl_scan(want, lane, &npair);
                        const int take = max(0, min(want, kTrussWave - s0));
                        npair = min(npair, kTrussWave);
                        for (int e = 0; e < take; ++e) {
                            const int k = __builtin_ctzll(hit);
                            hit &= hit - 1;
                            pairs[s0 + e] = (lane << 6) | k;
                        }
                        __syncthreads();   // the pairs are written
                        const int pr = lane < npair ? pairs[lane] : 0;
                        const int o = pr >> 6, k = pr & 63;
                        // the owner's query, from its lane (every lane takes part in the shuffles)
                        const float ox_ = __shfl(px, o, kWave), oy_ = __shfl(py, o, kWave);
                        const float oz_ = __shfl(pz, o, kWave), oh = __shfl(hi, o, kWave);
                        const int oid = __shfl(myid, o, kWave), otl = __shfl(ti_int | (li << 1), o, kWave);
                        const int os0 = __shfl(s0, o, kWave), ocnt = __shfl(cnt, o, kWave);
                        const uint64_t okth = __shfl(kth, o, kWave);
                        uint64_t key = 0;
                        bool valid = false;
                        if (lane < npair) {
                            const float4 c = cpos[k];
                            const int2 mt = cms[k];
                            const float ex = c.x - ox_, ey = c.y - oy_, ez = c.z - oz_;
                            const float d2 = ex * ex + ey * ey + ez * ez;
                            const float rj = half * (oh + c.w);
                            valid = d2 < rj * rj;
                            // not the node itself; no interior bar across labels
                            valid = valid && mt.x != oid &&
                                    !((otl & 1) && (mt.y & 0xFF) == TN_INTERIOR && (mt.y >> 8) != (otl >> 1));
                            if (valid && !H.prune) {   // the reference's own scan region only
                                const int L = r_lev[crow[k]];
                                const float rr = 0.5f * (tskin * oh + H.b0 * (float)(1 << L));
                                const int cx = cell_coord(H, L, c.x, H.ox, 0);
                                valid = cx >= cell_coord(H, L, ox_ - rr, H.ox, 0) &&
                                        cx <= cell_coord(H, L, ox_ + rr, H.ox, 0);
                            }
                            key = list_key(sqrtf(d2), mt.x);
                            valid = valid && key < okth;
                        }
                        const uint64_t vm = __ballot(valid);
                        // the owner's pairs are [s0, s0 + take) of its round: a valid pair's place is
                        // the owner's count plus the valid pairs of the owner before it
                        if (valid) {
                            const int at = ocnt + __popcll(vm & lanemask_lt() & ~((1ull << (os0 & 63)) - 1));
                            cb[at * kTrussWave + o] = (uint32_t)cslot[k];
                        }
                        if (take > 0) {
                            const uint64_t span = take == kTrussWave ? ~0ull : ((1ull << take) - 1);
                            cnt += __popcll((vm >> s0) & span);
                        }
                        __syncthreads();   // the pairs are consumed
                    }
                }
            };
            const int lmin = wave_min(ingroup ? max(Li - 1, 0) : INT_MAX);
            const int lmax = wave_max(ingroup ? min(Li + 1, H.nlev - 1) : -1);
            for (int L = lmin; L <= lmax; ++L) {
                const bool mine = ingroup && L >= Li - 1 && L <= Li + 1;
                const float bL = H.b0 * (float)(1 << L);
                const float r = 0.5f * (Ri + bL);
                int x0 = INT_MAX, x1 = -1, y0 = INT_MAX, y1 = -1, z0 = INT_MAX, z1 = -1;
                if (mine) {
                    x0 = cell_coord(H, L, px - r, H.ox, 0);
                    x1 = cell_coord(H, L, px + r, H.ox, 0);
                    y0 = cell_coord(H, L, py - r, H.oy, 1);
                    y1 = cell_coord(H, L, py + r, H.oy, 1);
                    z0 = cell_coord(H, L, pz - r, H.oz, 2);
                    z1 = cell_coord(H, L, pz + r, H.oz, 2);
                }
                const int Y0 = wave_min(y0), Y1 = wave_max(y1), Z0 = wave_min(z0), Z1 = wave_max(z1);
                // rows beyond reach are skipped (a valid bar is shorter than r; margin for rounding)
                const float rw = r * 1.0001f + 1e-3f * bL + 2e-6f * (fabsf(px) + fabsf(py) + fabsf(pz));
                const float rr = rw * rw;
                // the level's grid (uniform: kept in scalar registers through the rows)
                const int nx = __builtin_amdgcn_readfirstlane(H.dim[L][0]);
                const int ny = __builtin_amdgcn_readfirstlane(H.dim[L][1]);
                const int nz1 = __builtin_amdgcn_readfirstlane(H.dim[L][2]) - 1, ny1 = ny - 1;
                const int koff = __builtin_amdgcn_readfirstlane(H.off[L]);
                const float inv_bL = 1.0f / bL;
                for (int cz = Z0; cz <= Z1; ++cz) {
                    const float zlo = cz == 0 ? -inf : H.oz + (float)cz * bL;
                    const float zhi = cz == nz1 ? inf : H.oz + (float)(cz + 1) * bL;
                    const float dz = fmaxf(0.0f, fmaxf(zlo - pz, pz - zhi));
                    for (int cy = Y0; cy <= Y1; ++cy) {
                        const float ylo = cy == 0 ? -inf : H.oy + (float)cy * bL;
                        const float yhi = cy == ny1 ? inf : H.oy + (float)(cy + 1) * bL;
                        const float dy = fmaxf(0.0f, fmaxf(ylo - py, py - yhi));
                        const bool row_own = mine && cz >= z0 && cz <= z1 && cy >= y0 && cy <= y1;
                        int xa = INT_MAX, xb = -1;
                        if (row_own) {
                            if (!H.prune) {
                                xa = x0;
                                xb = x1;
                            } else {
                                // The x-extent of the ball of radius rw in this row. A valid
                                // candidate lies at least (rw - rj) / 2 >= 5e-4 bL inside it (rj <= r,
                                // the bar's reach), far more than the error of the hardware square
                                // root and of the multiplication by 1 / bL (below 1e-4 cells), so
                                // these approximate cells never drop a candidate.
                                const float rem = rr - dz * dz - dy * dy;
                                if (rem >= 0.0f) {
                                    const float w = __builtin_amdgcn_sqrtf(rem);
                                    const int ca = (int)floorf((px - w - H.ox) * inv_bL);
                                    const int cb_ = (int)floorf((px + w - H.ox) * inv_bL);
                                    xa = max(x0, min(max(ca, 0), nx - 1));
                                    xb = min(x1, min(max(cb_, 0), nx - 1));
                                }
                            }
                        }
                        const int XA = wave_min(xa), XB = wave_max(xb);
                        if (XA > XB) {
                            continue;
                        }
                        const uint64_t scan = __ballot(xa <= xb);   // the lanes with cells in the row
                        if (lane == 0) {
                            const int krow = koff + nx * (cy + ny * cz);
                            r_key0[nr] = krow + XA;
                            r_key1[nr] = krow + XB + 1;
                            r_mask[nr] = scan;
                            r_lev[nr] = L;
                        }
                        if (++nr == kMaxRanges) {
                            flush();
                        }
                    }
                }
            }
            if (nr > 0) {
                flush();
            }
        }   // the groups
        if (active) {   // (distance, id) order; the list holds slots
            uint64_t kk[32];
#pragma unroll
            for (int e = 0; e < 32; ++e) {
                kk[e] = e < cnt ? slot_key(in.Ps, in.Ms, (int)cb[e * kTrussWave + lane], px, py, pz) : ~0ull;
            }
            bitonic32(kk);
            const int m = cnt < kK ? cnt : kK;
            nnb[q] = m;
#pragma unroll
            for (int k = 0; k < kK; ++k) {
                if (k < m) {
                    nbrT[(size_t)k * n + q] = in.slot_of[(int)(uint32_t)kk[k]];
                }
            }
        }
        __syncthreads();   // the buffers are reused by the next block of queries
    }
}

// ---- own-list refresh: one wave per node --------------------------------------------
//
// The nodes past a full skin refresh their own list every iteration
// (src/v2m_particles.cpp:925-931: the current position and size against the cells
// and candidate positions of the last rebuild; then P0 = P). They are few (about
// 0.1% of the nodes) and scattered, so each gets a whole wave: the lanes test the
// node's candidates 64 at a time with coalesced loads, valid ones are appended to
// the wave's list in LDS, and the K best in (distance, id) order are picked by rank
// (the order is total, so ranks are distinct). The scan region and the row pruning
// are k_truss's for a single lane.

constexpr int kRefreshCap = 256;   // LDS list entries (cut back to the best K when full)

// Keep the K best of the cnt entries of (cd, cj), in order, at [0, min(cnt, K)).
__device__ __forceinline__ int keep_best(float* cd, int* cj, int cnt, int lane) {
    constexpr int kPer = kRefreshCap / kWave;
    float d[kPer];
    int j[kPer], rank[kPer];
#pragma unroll
    for (int t = 0; t < kPer; ++t) {
        const int e = lane + t * kWave;
        d[t] = e < cnt ? cd[e] : 0.0f;
        j[t] = e < cnt ? cj[e] : 0;
        rank[t] = 0;
    }
    for (int e = 0; e < cnt; ++e) {
        const float de = cd[e];
        const int je = cj[e];
#pragma unroll
        for (int t = 0; t < kPer; ++t) {
            rank[t] += before(de, je, d[t], j[t]) ? 1 : 0;
        }
    }
    __syncthreads();   // every entry is read before any is moved
#pragma unroll
    for (int t = 0; t < kPer; ++t) {
        if (lane + t * kWave < cnt && rank[t] < kK) {
            cd[rank[t]] = d[t];
            cj[rank[t]] = j[t];
        }
    }
    __syncthreads();
    return cnt < kK ? cnt : kK;
}

__device__ __forceinline__ void k_truss_node_body(Hash H, TrussIn in, int n, const int32_t* __restrict__ list, const Ctl* __restrict__ ctl, int32_t* __restrict__ nbrT, int32_t* __restrict__ nnb, float4* __restrict__ P0, int bid, int nblk) {
    __shared__ float cd[kRefreshCap];
    __shared__ int cj[kRefreshCap];
    const int lane = threadIdx.x;
    const int m = ctl->nrefresh;
    const float tskin = in.t + in.skin;
    for (int q = bid; q < m; q += nblk) {
        const int i = list[q];   // the node's slot
        const float4 xi = in.X[i];
        const float px = xi.x, py = xi.y, pz = xi.z, hi = xi.w;
        const float Ri = tskin * hi;
        const int Li = level_for(H, Ri);
        const int ti_int = in.typ[i] == TN_INTERIOR;
        const int li = in.lab[i];
        int cnt = 0;   // the wave's list length (the same in every lane)
        const int l0 = Li - 1 > 0 ? Li - 1 : 0, l1 = Li + 1 < H.nlev - 1 ? Li + 1 : H.nlev - 1;
        for (int L = l0; L <= l1; ++L) {
            const float bL = H.b0 * (float)(1 << L);
            const float r = 0.5f * (Ri + bL);
            const int x0 = cell_coord(H, L, px - r, H.ox, 0), x1 = cell_coord(H, L, px + r, H.ox, 0);
            const int y0 = cell_coord(H, L, py - r, H.oy, 1), y1 = cell_coord(H, L, py + r, H.oy, 1);
            const int z0 = cell_coord(H, L, pz - r, H.oz, 2), z1 = cell_coord(H, L, pz + r, H.oz, 2);
            const float rw = r * 1.0001f + 1e-3f * bL + 2e-6f * (fabsf(px) + fabsf(py) + fabsf(pz));
            const float rr = rw * rw;
            const int ny1 = H.dim[L][1] - 1, nz1 = H.dim[L][2] - 1;
            for (int cz = z0; cz <= z1; ++cz) {
                const float zlo = cz == 0 ? -__builtin_huge_valf() : H.oz + (float)cz * bL;
                const float zhi = cz == nz1 ? __builtin_huge_valf() : H.oz + (float)(cz + 1) * bL;
                const float dz = fmaxf(0.0f, fmaxf(zlo - pz, pz - zhi));
                for (int cy = y0; cy <= y1; ++cy) {
                    const float ylo = cy == 0 ? -__builtin_huge_valf() : H.oy + (float)cy * bL;
                    const float yhi = cy == ny1 ? __builtin_huge_valf() : H.oy + (float)(cy + 1) * bL;
                    const float dy = fmaxf(0.0f, fmaxf(ylo - py, py - yhi));
                    int xa = x0, xb = x1;
                    if (H.prune) {
                        const float rem = rr - dz * dz - dy * dy;
                        if (rem < 0.0f) {
                            continue;
                        }
                        const float wx = sqrtf(rem);
                        xa = max(x0, cell_coord(H, L, px - wx, H.ox, 0));
                        xb = min(x1, cell_coord(H, L, px + wx, H.ox, 0));
                        if (xa > xb) {
                            continue;
                        }
                    }
                    const int krow = H.off[L] + H.dim[L][0] * (cy + H.dim[L][1] * cz);
                    const int sa = in.cstart[krow + xa], sb = in.cstart[krow + xb + 1];
                    for (int base = sa; base < sb; base += kWave) {
                        const int s = base + lane;
                        bool ok = false;
                        float dist = 0.0f;
                        int j = 0;
                        if (s < sb) {
                            const float4 c = in.Ps[s];
                            const float ex = c.x - px, ey = c.y - py, ez = c.z - pz;
                            const float d2 = ex * ex + ey * ey + ez * ez;
                            const float rj = tskin * 0.5f * (hi + c.w);
                            if (d2 < rj * rj) {   // slot s: the candidate's current type, its id
                                ok = s != i && !(ti_int && in.typ[s] == TN_INTERIOR && in.lab[s] != li);
                                j = in.Ms[s].x;
                                dist = sqrtf(d2);
                            }
                        }
                        const uint64_t vm = __ballot(ok);
                        if (vm) {
                            const int add = __popcll(vm);
                            if (cnt + add > kRefreshCap) {
                                cnt = keep_best(cd, cj, cnt, lane);
                            }
                            if (ok) {
                                const int pos = cnt + __popcll(vm & lanemask_lt());
                                cd[pos] = dist;
                                cj[pos] = j;
                            }
                            cnt += add;
                        }
                    }
                }
            }
        }
        __syncthreads();   // the appends are visible to every lane
        const int kk = keep_best(cd, cj, cnt, lane);
        if (lane < kk) {
            nbrT[(size_t)lane * n + i] = in.slot_of[cj[lane]];
        }
        if (lane == 0) {
            nnb[i] = kk;
            P0[i] = xi;
        }
        __syncthreads();   // the list is read before the next node reuses it
    }
}

__global__ __launch_bounds__(kWave) void k_truss_node(Hash H, TrussIn in, int n, const int32_t* __restrict__ list, const Ctl* __restrict__ ctl, int32_t* __restrict__ nbrT, int32_t* __restrict__ nnb, float4* __restrict__ P0) {
    k_truss_node_body(H, in, n, list, ctl, nbrT, nnb, P0, blockIdx.x, gridDim.x);
}

// the slots of the nodes flagged for an own-list refresh, appended block by block
// (in slot order, which is cell order as of the last rebuild, within a block)
__device__ __forceinline__ void k_refresh_list_body(const uint8_t* __restrict__ flag, int n, Ctl* __restrict__ ctl, int32_t* __restrict__ list, int bid, int nblk) {
    __shared__ int woff[kBlock / kWave];
    __shared__ int base;
    // k_control has run: a rebuild (which rebuilds every list and sets P0 = P) or the
    // end of the loop follows, so no refreshed list would ever be read
    if (ctl->rebuild || ctl->done) {
        return;
    }
    const int s = bid * kBlock + threadIdx.x;
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x >> 6;
    const bool f = s < n && flag[s];
    const uint64_t mask = __ballot(f);
    if (lane == 0) {
        woff[w] = __popcll(mask);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int tot = 0;
        for (int k = 0; k < kBlock / kWave; ++k) {
            const int c = woff[k];
            woff[k] = tot;
            tot += c;
        }
        base = tot ? atomicAdd(&ctl->nrefresh, tot) : 0;
    }
    __syncthreads();
    if (f) {
        list[base + woff[w] + __popcll(mask & lanemask_lt())] = s;
    }
}

__global__ __launch_bounds__(kBlock) void k_refresh_list(const uint8_t* __restrict__ flag, int n, Ctl* __restrict__ ctl, int32_t* __restrict__ list) {
    k_refresh_list_body(flag, n, ctl, list, blockIdx.x, gridDim.x);
}

// ---- node counts by type: interface, junction, interior, corner ------------------------

__device__ __forceinline__ int type_rank(int t) {
    return t == TN_INTERFACE ? 0 : (t == TN_JUNCTION ? 1 : (t == TN_INTERIOR ? 2 : 3));
}

// lanes of this wave whose node has type rank r (4 ballots)
__device__ __forceinline__ void rank_masks(bool valid, int r, uint64_t* m) {
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        m[k] = __ballot(valid && r == k);
    }
}

__global__ __launch_bounds__(kBlock) void k_order_count(const uint8_t* __restrict__ typ, int n, Ctl* __restrict__ ctl) {
    __shared__ int c[4];
    if (threadIdx.x < 4) {
        c[threadIdx.x] = 0;
    }
    __syncthreads();
    const int i = blockIdx.x * kBlock + threadIdx.x;
    const bool valid = i < n;
    uint64_t m[4];
    rank_masks(valid, valid ? type_rank(typ[i] & 3) : 0, m);
    if ((threadIdx.x & (kWave - 1)) == 0) {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (m[k]) {
                atomicAdd(&c[k], __popcll(m[k]));
            }
        }
    }
    __syncthreads();
    if (threadIdx.x < 4 && c[threadIdx.x]) {
        atomicAdd(&ctl->order_cnt[threadIdx.x], c[threadIdx.x]);
    }
}

// ---- force --------------------------------------------------------------------------

// append i to queue for the lanes with flag set (one atomic on *count per wave)
__device__ __forceinline__ void wave_append(bool flag, int i, int32_t* count, int32_t* __restrict__ queue) {
    const uint64_t m = __ballot(flag);
    if (m) {
        const int lead = __ffsll((long long)m) - 1;
        int base = 0;
        if ((int)(threadIdx.x & (kWave - 1)) == lead) {
            base = atomicAdd(count, __popcll(m));
        }
        base = __shfl(base, lead, kWave);
        if (flag) {
            queue[base + __popcll(m & lanemask_lt())] = i;
        }
    }
}


// The forces; with queue_rest, also the junction nodes queued for the general move
// pass (their moves need no result of k_move_easy, so that pass can start as soon as
// the forces are done) and flagged in skip for k_move_easy, which leaves them alone
// without reading their type (the general pass may be rewriting it).
// The forces, and the queue of the general move pass that starts beside k_move_easy:
// the junction nodes and, when `sticky` is given, the nodes a lean pass has handed back
// before (their move is the same whichever pass computes it; queued here, it leaves the
// iteration's critical path).
__device__ __forceinline__ void k_force_body(const float4* __restrict__ X, const uint8_t* __restrict__ typ, const int32_t* __restrict__ nbrT, const int32_t* __restrict__ nnb, float fscale, float fsurf, int n, float4* __restrict__ F, Ctl* __restrict__ ctl, int32_t* __restrict__ queue_rest, uint8_t* __restrict__ skip, const uint8_t* __restrict__ sticky, int bid, int nblk) {
    const int i = bid * kBlock + threadIdx.x;
    bool junction = false;
    if (i < n) {
        F[i] = bar_force(X, typ, nbrT, nnb, fscale, fsurf, i, n);
        if (queue_rest) {
            junction = typ[i] == TN_JUNCTION || (sticky && sticky[i]);
            skip[i] = junction ? 1 : 0;
        }
    }
    if (queue_rest) {
        wave_append(junction, i, &ctl->nmove_rest, queue_rest);
    }
}

__global__ __launch_bounds__(kBlock) void k_force(const float4* __restrict__ X, const uint8_t* __restrict__ typ, const int32_t* __restrict__ nbrT, const int32_t* __restrict__ nnb, float fscale, float fsurf, int n, float4* __restrict__ F, Ctl* __restrict__ ctl, int32_t* __restrict__ queue_rest, uint8_t* __restrict__ skip, const uint8_t* __restrict__ sticky) {
    k_force_body(X, typ, nbrT, nnb, fscale, fsurf, n, F, ctl, queue_rest, skip, sticky, blockIdx.x, gridDim.x);
}

// ---- move (v2m_move, src/opencl/v2m_particle_body.cl:844-935) ----------------------

struct MoveArgs {
    float dt, maxstep, snap;
    int fire;
    int lanes;   // nodes per wave of the general pass at least (a launch shape; results do not depend on it)
};

__device__ __forceinline__ uint32_t pack_part(int b, int c) {
    return (uint32_t)(b & 0xFFFF) | ((uint32_t)(c & 0xFFFF) << 16);
}

// The FIRE / Jacobi part of v2m_move (src/opencl/v2m_particle_body.cl:853-912) for a
// node that is not a corner: the step s, capped at maxstep h, the power a . v, and
// whether the node moves at all (a settled node, |s| < 5e-4 h, stays put).
struct NodeStep {
    float s[3];
    float pw;
    bool moves;
};

// (v_at(): the node's velocity, read only when FIRE uses it)
template <class VAt>
__device__ __forceinline__ NodeStep node_step_with(const MoveArgs& ma, const Ctl* __restrict__ ctl, float4 Fi,
                                                   VAt v_at, float h) {
    NodeStep r;
    r.pw = 0.0f;
    const float stiff = Fi.w > 1.0f ? Fi.w : 1.0f;
    const float ac[3] = { Fi.x / stiff, Fi.y / stiff, Fi.z / stiff };
    if (ma.fire) {
        const float fdt = ctl->dt, falpha = ctl->alpha;
        float v[3] = { 0.0f, 0.0f, 0.0f };
        if (!ctl->zero_v) {
            const float4 v4 = v_at();
            v[0] = v4.x;
            v[1] = v4.y;
            v[2] = v4.z;
        }
        const float pv = ac[0] * v[0] + ac[1] * v[1] + ac[2] * v[2];
        r.pw = pv;
        if (pv < 0.0f) {
            v[0] = v[1] = v[2] = 0.0f;
        } else {
            const float vn = sqrtf(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
            const float an = sqrtf(ac[0] * ac[0] + ac[1] * ac[1] + ac[2] * ac[2]);
            if (an > 0.0f) {
                for (int k = 0; k < 3; ++k) {
                    v[k] = (1.0f - falpha) * v[k] + falpha * vn * ac[k] / an;
                }
            }
        }
        for (int k = 0; k < 3; ++k) {
            v[k] += fdt * ac[k];
            r.s[k] = fdt * v[k];
        }
    } else {
        for (int k = 0; k < 3; ++k) {
            r.s[k] = ma.dt * ac[k];
        }
    }
    const float sl = sqrtf(r.s[0] * r.s[0] + r.s[1] * r.s[1] + r.s[2] * r.s[2]);
    r.moves = !(sl < 5e-4f * h);
    if (r.moves && sl > ma.maxstep * h) {
        for (int k = 0; k < 3; ++k) {
            r.s[k] *= ma.maxstep * h / sl;
        }
    }
    return r;
}

__device__ __forceinline__ NodeStep node_step(const MoveArgs& ma, const Ctl* __restrict__ ctl, float4 Fi,
                                              const float4* __restrict__ V, int i, float h) {
    return node_step_with(ma, ctl, Fi, [&] { return V[i]; }, h);
}

// the writes that end v2m_move for node i (src/opencl/v2m_particle_body.cl:921-934)
__device__ __forceinline__ void move_store(const MoveArgs& ma, const Ctl* __restrict__ ctl, int i, float4 x, int ty0,
                                           uint32_t pp, const NodeState& ns, float m, float p_w,
                                           float4* __restrict__ X, float4* __restrict__ V,
                                           uint8_t* __restrict__ typ, uint32_t* __restrict__ part,
                                           float* __restrict__ mv, float* __restrict__ pw) {
    if (m > 0.0f) {
        X[i] = make_float4(ns.p.x, ns.p.y, ns.p.z, ns.hn);
    }
    if (ns.ty != ty0) {
        typ[i] = (uint8_t)ns.ty;
    }
    const uint32_t np = pack_part(ns.b, ns.c);
    if (np != pp) {
        part[i] = np;
    }
    if (ma.fire) {
        const float fdt = ctl->dt;
        float4 vout = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (m > 0.0f && ns.ty == ty0) {
            vout = make_float4((ns.p.x - x.x) / fdt, (ns.p.y - x.y) / fdt, (ns.p.z - x.z) / fdt, 0.0f);
        }
        V[i] = vout;
        pw[i] = p_w;
    }
    mv[i] = m;
}

// Pass 1 of the move, over every node in slot order but the junction nodes (queued
// by k_force for the general pass): corners, settled nodes and interior nodes whose
// step stays in their own region with no interface within snap range are finished
// here, exactly as v2m_move finishes them. The interface nodes and the interior
// nodes that cross or may snap are queued, untouched, for the lean k_move passes.
// Most nodes are interior, and this pass needs a fraction of the registers of the
// full trapped step.
// Append i to queue a (lanes with fa set) and to queue b (fb), one atomic per block
// and queue: the 2 M-node pass would otherwise send tens of thousands of atomics to
// the same two counters. Every thread of the block must arrive here.
__device__ __forceinline__ void block_append2(bool fa, bool fb, int i, int32_t* ca, int32_t* __restrict__ qa,
                                              int32_t* cb, int32_t* __restrict__ qb) {
    constexpr int kW = kBlock / kWave;
    __shared__ int cnt[2][kW + 1];
    const uint64_t ma = __ballot(fa), mb = __ballot(fb);
    const int w = threadIdx.x >> 6, lane = threadIdx.x & (kWave - 1);
    if (lane == 0) {
        cnt[0][w] = __popcll(ma);
        cnt[1][w] = __popcll(mb);
    }
    __syncthreads();
    if (threadIdx.x < 2) {   // the block's offsets of its waves, then one atomic per queue
        int tot = 0;
        for (int k = 0; k < kW; ++k) {
            const int c = cnt[threadIdx.x][k];
            cnt[threadIdx.x][k] = tot;
            tot += c;
        }
        cnt[threadIdx.x][kW] = tot ? atomicAdd(threadIdx.x == 0 ? ca : cb, tot) : 0;
    }
    __syncthreads();
    const uint64_t below = lanemask_lt();
    if (fa) {
        qa[cnt[0][kW] + cnt[0][w] + __popcll(ma & below)] = i;
    }
    if (fb) {
        qb[cnt[1][kW] + cnt[1][w] + __popcll(mb & below)] = i;
    }
}

// kPlain: every grid of the launch is a label volume (gm == 0). The compiler then
// drops the gray-scale and shape-field branches of the field queries, and the hot
// kernels keep their register budget; other grids take the kPlain = false variants.
template <bool kPlain>
__device__ __forceinline__ void k_move_easy_body(Field f, MoveArgs ma, Ctl* __restrict__ ctl, int n, const float4* __restrict__ F, float4* __restrict__ X, float4* __restrict__ V, uint8_t* __restrict__ typ, uint32_t* __restrict__ part, const uint16_t* __restrict__ lab, float* __restrict__ mv, float* __restrict__ pw, int32_t* __restrict__ queue, int32_t* __restrict__ queue_int, const uint8_t* __restrict__ skip, int bid, int nblk) {
    if (kPlain) {
        __builtin_assume(f.gm == 0);
    }
    const int g = bid * kBlock + threadIdx.x;
    int i = 0;
    bool surf = false, inner = false;   // the two lean queues
    // The node's own values, read at once rather than each behind the test that needs
    // it (a node another pass moves reads values it then leaves unused).
    const bool in = g < n;
    const int sk = in ? skip[g] : 1;
    const int ty0 = in ? typ[g] : 0;
    const float4 x = in ? X[g] : make_float4(0.0f, 0.0f, 0.0f, 1.0f);
    const float4 Fi = in ? F[g] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    const float4 Vi = in && ma.fire ? V[g] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    const int a = in ? lab[g] : 0;
    const uint32_t pp = in ? part[g] : 0u;
    if (in && !sk) {   // (junction nodes: queued by k_force for the general pass)
        i = g;
        if (ty0 == TN_INTERFACE) {
            surf = true;
        } else if (ty0 == TN_CORNER) {
            if (ma.fire) {
                V[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                pw[i] = 0.0f;
            }
            mv[i] = 0.0f;
        } else {
            const float h = x.w;
            const NodeStep st = node_step_with(ma, ctl, Fi, [&] { return Vi; }, h);
            if (!st.moves) {
                if (ma.fire) {
                    V[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                    pw[i] = st.pw;
                }
                mv[i] = 0.0f;
            } else {
                // trapped_step's interior branch up to its first decision
                const P3 p = p3(x.x, x.y, x.z);
                const P3 q = p3(p.x + st.s[0], p.y + st.s[1], p.z + st.s[2]);
                // h at the target, loaded before the owner query's chain of loads (it is
                // needed unless the node turns out to cross or to snap)
                const float hq = size_at(f, q.x, q.y, q.z);
                const OwnerPsi own = owner_psi_at(f, a, q);
                if (own.best != a) {
                    inner = true;
                } else if (own.sec != kNoLab && ma.snap > 0.0f) {
                    const float4 gq = own.g;
                    const float gn = sqrtf(gq.x * gq.x + gq.y * gq.y + gq.z * gq.z);
                    inner = gn > 1e-9f && gq.w / gn < ma.snap * h;
                }
                if (!inner) {   // the rest of trapped_step for an interior node that stays interior
                    NodeState ns;
                    ns.p = p;
                    ns.hn = h;
                    ns.ty = TN_INTERIOR;
                    ns.b = (int)(pp & 0xFFFF);
                    ns.c = (int)(pp >> 16);
                    float m = 0.0f;
                    const float lim = (ma.maxstep + fmaxf(ma.snap, 0.5f)) * h;
                    const float m2 = dist2(q, p);
                    if (!(m2 <= lim * lim)) {
                        ns.b = kNoLab;
                    } else {
                        ns.p = q;
                        ns.hn = hq;
                        m = sqrtf(m2) / h;
                    }
                    move_store(ma, ctl, i, x, TN_INTERIOR, pp, ns, m, st.pw, X, V, typ, part, mv, pw);
                }
            }
        }
    }
    // queue the rest, one atomic per block and queue
    block_append2(surf, inner, i, &ctl->nmove, queue, &ctl->nmove_int, queue_int);
}

template <bool kPlain>
__global__ __launch_bounds__(kBlock) void k_move_easy(Field f, MoveArgs ma, Ctl* __restrict__ ctl, int n, const float4* __restrict__ F, float4* __restrict__ X, float4* __restrict__ V, uint8_t* __restrict__ typ, uint32_t* __restrict__ part, const uint16_t* __restrict__ lab, float* __restrict__ mv, float* __restrict__ pw, int32_t* __restrict__ queue, int32_t* __restrict__ queue_int, const uint8_t* __restrict__ skip) {
    k_move_easy_body<kPlain>(f, ma, ctl, n, F, X, V, typ, part, lab, mv, pw, queue, queue_int, skip, blockIdx.x,
                             gridDim.x);
}

#ifndef TN_REST_LANES
#define TN_REST_LANES 4
#endif
constexpr int kRestLanes = TN_REST_LANES;   // nodes per wave in k_move<kGeneral>
// a lean pass spreads up to this many nodes per wave before it fills whole waves
#ifndef TN_LEAN_SPREAD
#define TN_LEAN_SPREAD 4
#endif
constexpr int kLeanSpread = TN_LEAN_SPREAD;

// waves per SIMD the full move kernel is compiled for (its register budget: 512 / waves)
#ifndef TN_MOVE_WAVES
#define TN_MOVE_WAVES 1
#endif

// v2m_move for queued node i; true: handed back (a lean mode, a third label near)
template <int kMode>
__device__ __forceinline__ bool move_node(const Field& f, const MoveArgs& ma, const Ctl* __restrict__ ctl, int i,
                                          const float4* __restrict__ F, float4* __restrict__ X,
                                          float4* __restrict__ V, uint8_t* __restrict__ typ,
                                          uint32_t* __restrict__ part, const uint16_t* __restrict__ lab,
                                          float* __restrict__ mv, float* __restrict__ pw) {
    const float4 x = X[i];
    const float h = x.w;   // h at the node: v2m_h_at(P_i), kept current by every move
    const int ty0 = kMode == kLeanInterface ? TN_INTERFACE : (kMode == kLeanInterior ? TN_INTERIOR : typ[i]);
    const uint32_t pp = part[i];
    NodeState ns;
    ns.p = p3(x.x, x.y, x.z);
    ns.hn = h;
    ns.ty = ty0;
    ns.b = (int)(pp & 0xFFFF);
    ns.c = (int)(pp >> 16);
    float m = 0.0f;
    const NodeStep st = node_step(ma, ctl, F[i], V, i, h);   // queued nodes are never corners
    if (st.moves) {
        m = trapped_step<kMode>(f, h, p3(st.s[0], st.s[1], st.s[2]), ma.maxstep, ma.snap, lab[i], ns);
        if (kMode != kGeneral && m == kDeferred) {
            return true;
        }
    }
    move_store(ma, ctl, i, x, ty0, pp, ns, m, st.pw, X, V, typ, part, mv, pw);
    return false;
}

// The whole move in voxel-trap mode (v2m_move with voxmode, src/opencl/v2m_particle_body.cl:
// 844-935 and 549-650): every node in one pass; corners and settled nodes stay.
// The voxel-trap move, pass 1, a lane per node: corners, settled nodes and interior
// nodes whose step meets no wall and no wall within snap range are finished here; the
// rest (face and edge scans) are queued, untouched, for pass 2.
__device__ __forceinline__ void k_move_voxel_body(const Field& f, const MoveArgs& ma, const Ctl* __restrict__ ctl,
                                                  int n, const float4* __restrict__ F, float4* __restrict__ X,
                                                  float4* __restrict__ V, uint8_t* __restrict__ typ,
                                                  uint32_t* __restrict__ part, const uint16_t* __restrict__ lab,
                                                  float* __restrict__ mv, float* __restrict__ pw,
                                                  int32_t* __restrict__ queue, int32_t* __restrict__ qcount, int bid) {
    const int i = bid * kBlock + threadIdx.x;
    bool defer = false;
    if (i < n) {
        const float4 x = X[i];
        const float h = x.w;
        const int ty0 = typ[i];
        const uint32_t pp = part[i];
        NodeState ns;
        ns.p = p3(x.x, x.y, x.z);
        ns.hn = h;
        ns.ty = ty0;
        ns.b = (int)(pp & 0xFFFF);
        ns.c = (int)(pp >> 16);
        float m = 0.0f, p_w = 0.0f;
        if (ty0 != TN_CORNER) {
            const NodeStep st = node_step(ma, ctl, F[i], V, i, h);
            p_w = st.pw;
            if (st.moves) {
                m = voxel_step<false>(f, h, ma.snap, lab[i], st.s, ns);
                defer = m == kDeferred;
            }
        }
        if (!defer) {
            move_store(ma, ctl, i, x, ty0, pp, ns, m, p_w, X, V, typ, part, mv, pw);
        }
    }
    wave_append(defer, i, qcount, queue);
}

__global__ __launch_bounds__(kBlock) void k_move_voxel(Field f, MoveArgs ma, const Ctl* __restrict__ ctl, int n,
                                                       const float4* __restrict__ F, float4* __restrict__ X,
                                                       float4* __restrict__ V, uint8_t* __restrict__ typ,
                                                       uint32_t* __restrict__ part, const uint16_t* __restrict__ lab,
                                                       float* __restrict__ mv, float* __restrict__ pw,
                                                       int32_t* __restrict__ queue, int32_t* __restrict__ qcount) {
    k_move_voxel_body(f, ma, ctl, n, F, X, V, typ, part, lab, mv, pw, queue, qcount, blockIdx.x);
}

// Pass 2: a wave per queued node, every lane on the same node (the same values in
// every lane), the face and edge scans split over the wave; lane 0 stores.
__device__ __forceinline__ void k_move_voxel_wave_body(const Field& f, const MoveArgs& ma, const Ctl* __restrict__ ctl,
                                                       const int32_t* __restrict__ queue,
                                                       const int32_t* __restrict__ qcount,
                                                       const float4* __restrict__ F, float4* __restrict__ X,
                                                       float4* __restrict__ V, uint8_t* __restrict__ typ,
                                                       uint32_t* __restrict__ part, const uint16_t* __restrict__ lab,
                                                       float* __restrict__ mv, float* __restrict__ pw, int bid,
                                                       int nblk) {
    constexpr int kWaves = kBlock / kWave;
    const int count = *qcount;
    for (int g = bid * kWaves + (int)(threadIdx.x >> 6); g < count; g += nblk * kWaves) {
        const int i = queue[g];
        const float4 x = X[i];
        const float h = x.w;
        const int ty0 = typ[i];
        const uint32_t pp = part[i];
        NodeState ns;
        ns.p = p3(x.x, x.y, x.z);
        ns.hn = h;
        ns.ty = ty0;
        ns.b = (int)(pp & 0xFFFF);
        ns.c = (int)(pp >> 16);
        const NodeStep st = node_step(ma, ctl, F[i], V, i, h);   // queued nodes move
        const float m = voxel_step<true>(f, h, ma.snap, lab[i], st.s, ns);
        if ((threadIdx.x & (kWave - 1)) == 0) {
            move_store(ma, ctl, i, x, ty0, pp, ns, m, st.pw, X, V, typ, part, mv, pw);
        }
    }
}

__global__ __launch_bounds__(kBlock) void k_move_voxel_wave(Field f, MoveArgs ma, const Ctl* __restrict__ ctl,
                                                            const int32_t* __restrict__ queue,
                                                            const int32_t* __restrict__ qcount,
                                                            const float4* __restrict__ F, float4* __restrict__ X,
                                                            float4* __restrict__ V, uint8_t* __restrict__ typ,
                                                            uint32_t* __restrict__ part,
                                                            const uint16_t* __restrict__ lab, float* __restrict__ mv,
                                                            float* __restrict__ pw) {
    k_move_voxel_wave_body(f, ma, ctl, queue, qcount, F, X, V, typ, part, lab, mv, pw, blockIdx.x, gridDim.x);
}

// Pass 2: v2m_move (src/opencl/v2m_particle_body.cl:844-935) for the queued nodes,
// *count of them. kLeanInterface / kLeanInterior: interface or interior nodes, with
// their branch of the trapped step only and no out-of-line calls (whose register
// budgets would bound the kernel's occupancy); a node with a third label near is
// appended to (defer_count, defer_queue) for a later kGeneral pass. kGeneral: the
// whole trapped step (junction nodes and the handed-back ones). Each node's move
// reads and writes only that node's state, so the passes over disjoint queues may
// run concurrently.
template <int kMode, bool kPlain>
__device__ __forceinline__ void k_move_body(Field f, MoveArgs ma, Ctl* __restrict__ ctl, const int32_t* __restrict__ queue, const float4* __restrict__ F, float4* __restrict__ X, float4* __restrict__ V, uint8_t* __restrict__ typ, uint32_t* __restrict__ part, const uint16_t* __restrict__ lab, float* __restrict__ mv, float* __restrict__ pw, const int32_t* __restrict__ qcount, int32_t* __restrict__ defer_count, int32_t* __restrict__ defer_queue, uint8_t* __restrict__ sticky, int bid, int nblk) {
    if (kPlain) {
        __builtin_assume(f.gm == 0);
    }
    // A fixed grid strides over the queue, whose length only the device knows. A
    // short queue's nodes follow long, divergent paths (projections onto interfaces
    // and junction curves), and a wave executes every path its lanes take, so they
    // are spread thinly over the grid's waves (lean passes: one to kLeanSpread nodes per
    // wave; the general pass: from kRestLanes, measured best); a long queue fills whole
    // waves, whose lanes then share the issue slots.
    constexpr bool kLean = kMode != kGeneral;
    constexpr int kWaves = kBlock / kWave;
    const int count = *qcount;
    const int waves = (int)nblk * kWaves;
    const int spread = (count + waves - 1) / waves;
    int lpw;
    if (kLean) {
        lpw = spread < 1 ? 1 : (spread > kLeanSpread ? kWave : spread);
    } else {
        const int rl = ma.lanes > 0 ? ma.lanes : kRestLanes;
        lpw = spread < rl ? rl : (spread > kWave ? kWave : spread);
    }
    const int per_block = lpw * kWaves;
    const int lane = threadIdx.x & (kWave - 1);
    const int slot = lane < lpw ? (threadIdx.x >> 6) * lpw + lane : -1;
    for (int base = bid * per_block; base < count; base += nblk * per_block) {
        const int g = base + slot;
        bool defer = false;
        int i = 0;
        if (slot >= 0 && g < count) {
            i = queue[g];
            defer = move_node<kMode>(f, ma, ctl, i, F, X, V, typ, part, lab, mv, pw);
        }
        if (kLean) {   // every lane of the wave is here
            wave_append(defer, i, defer_count, defer_queue);
            if (defer && sticky) {   // from the next iteration on, queued by k_force
                sticky[i] = 1;
            }
        }
    }
}

template <int kMode, bool kPlain>
__global__ __launch_bounds__(kBlock, TN_MOVE_WAVES) void k_move(Field f, MoveArgs ma, Ctl* __restrict__ ctl, const int32_t* __restrict__ queue, const float4* __restrict__ F, float4* __restrict__ X, float4* __restrict__ V, uint8_t* __restrict__ typ, uint32_t* __restrict__ part, const uint16_t* __restrict__ lab, float* __restrict__ mv, float* __restrict__ pw, const int32_t* __restrict__ qcount, int32_t* __restrict__ defer_count, int32_t* __restrict__ defer_queue, uint8_t* __restrict__ sticky) {
    k_move_body<kMode, kPlain>(f, ma, ctl, queue, F, X, V, typ, part, lab, mv, pw, qcount, defer_count, defer_queue,
                               sticky, blockIdx.x, gridDim.x);
}

constexpr int kMoveGrid = 2048;       // blocks of a lean k_move at most
constexpr int kMoveRestGrid = 512;    // blocks of k_move<kGeneral>: the waves resident at its register budget


// ---- per-iteration reductions + Verlet bookkeeping ----------------------------------

// The histogram bin of src/v2m_particles.cpp:881: clamp(20 + floor(log2f(m)), 0, 31).
// floor(log2f(m)) is the binary exponent of m, except just below a power of two,
// where the rounded log2f reaches the next integer; thr[e + 21] is the smallest
// float below 2^e whose log2f rounds up to e (from the host's log2f).
__device__ __forceinline__ int hist_bin(float m, const float* __restrict__ thr) {
    if (!(m > 0.0f)) {
        return 0;
    }
    int e;
    frexpf(m, &e);   // m in [2^(e-1), 2^e)
    int fl = e - 1;
    if (e >= -21 && e <= 13 && m >= thr[e + 21]) {
        fl = e;
    }
    const int b = 20 + fl;
    return b < 0 ? 0 : (b > 31 ? 31 : b);
}

constexpr int kHistThr = 35;   // exponents -21..13
constexpr int kStatsBlocks = 2048;   // k_stats' grid at most (8 blocks per CU)

// Per iteration: the histogram, the largest move, the count past skin/2, the
// flags of the nodes past a full skin (their lists are refreshed), and FIRE's
// power partial sums (src/v2m_particles.cpp:849-937). Counts are reduced per
// wave with ballots before any atomic.
__device__ __forceinline__ void k_stats_body(const float4* __restrict__ X, const float4* __restrict__ P0, const float* __restrict__ mv, const float* __restrict__ pw, int n, float skin, int fire, const float* __restrict__ thr, Ctl* __restrict__ ctl, uint8_t* __restrict__ refresh, double2* __restrict__ partials, int bid, int nblk) {
    __shared__ uint32_t lh[32];
    __shared__ uint32_t lmax;
    __shared__ int lhalf;
    __shared__ double lps[kBlock / kWave][2];
    const int tid = threadIdx.x, lane = tid & (kWave - 1);
    if (tid < 32) {
        lh[tid] = 0;
    }
    if (tid == 0) {
        lmax = 0;
        lhalf = 0;
    }
    double bs = 0.0, ba = 0.0;   // this block's power partial sums (thread 0), chunk by chunk
    // chunks of kBlock nodes, grid-strided: each block adds its counts to the control
    // block once, and its partial sums have a fixed shape
    for (int c = bid; (int64_t)c * kBlock < n; c += nblk) {
        __syncthreads();   // lps of the previous chunk has been read
        const int i = c * kBlock + tid;
        double ps = 0.0, pa = 0.0;
        int bin = -1;
        uint32_t mbits = 0;
        bool half = false;
        if (i < n) {
            const float m = mv[i];
            bin = hist_bin(m, thr);
            mbits = __float_as_uint(m);
            const float4 x = X[i], x0 = P0[i];
            const float ex = x.x - x0.x, ey = x.y - x0.y, ez = x.z - x0.z;
            const float h = x.w;
            const float m2 = ex * ex + ey * ey + ez * ez, s2 = skin * skin * h * h;
            const bool far = m2 > s2;
            refresh[i] = far ? 1 : 0;
            half = !far && m2 > 0.25f * s2;
            if (fire) {
                ps = (double)pw[i];
                pa = fabs(ps);
            }
        }
        // histogram: one LDS add per (wave, occupied bin)
        uint64_t left = __ballot(bin >= 0);
        while (left) {
            const int b = __shfl(bin, __ffsll((long long)left) - 1, kWave);
            const uint64_t same = __ballot(bin == b);
            if (lane == 0) {
                atomicAdd(&lh[b], (uint32_t)__popcll(same));
            }
            left &= ~same;
        }
#pragma unroll
        for (int d = kWave / 2; d > 0; d >>= 1) {
            mbits = max(mbits, (uint32_t)__shfl_xor((int)mbits, d, kWave));
        }
        const int nh = __popcll(__ballot(half));
        if (lane == 0) {
            atomicMax(&lmax, mbits);
            if (nh) {
                atomicAdd(&lhalf, nh);
            }
        }
        if (fire) {   // fixed-shape tree: the partial sums are reproducible
#pragma unroll
            for (int d = kWave / 2; d > 0; d >>= 1) {
                ps += __shfl_xor(ps, d, kWave);
                pa += __shfl_xor(pa, d, kWave);
            }
            if (lane == 0) {
                lps[tid >> 6][0] = ps;
                lps[tid >> 6][1] = pa;
            }
            __syncthreads();
            if (tid == 0) {
                bs += (lps[0][0] + lps[1][0]) + (lps[2][0] + lps[3][0]);
                ba += (lps[0][1] + lps[1][1]) + (lps[2][1] + lps[3][1]);
            }
        }
    }
    __syncthreads();
    if (tid < 32 && lh[tid]) {
        atomicAdd(&ctl->hist[tid], lh[tid]);
    }
    if (tid == 0) {
        atomicMax(&ctl->mmax_bits, lmax);
        if (lhalf) {
            atomicAdd(&ctl->nhalf, lhalf);
        }
        if (fire) {
            partials[bid] = make_double2(bs, ba);
        }
    }
}

__global__ __launch_bounds__(kBlock) void k_stats(const float4* __restrict__ X, const float4* __restrict__ P0, const float* __restrict__ mv, const float* __restrict__ pw, int n, float skin, int fire, const float* __restrict__ thr, Ctl* __restrict__ ctl, uint8_t* __restrict__ refresh, double2* __restrict__ partials) {
    k_stats_body(X, P0, mv, pw, n, skin, fire, thr, ctl, refresh, partials, blockIdx.x, gridDim.x);
}

// ---- field queries at given points (the queries the move makes, for checking) ---------

// phi_a (v2m_phi_f), psi_ab and its gradient (v2m_psi_grad), the owner and the
// runner-up (v2m_label_of) at X[3k..3k+2]; any output may be null.
__global__ __launch_bounds__(kBlock) void k_field_points(Field f, int a, int b, int m, const float* __restrict__ X,
                                                         float* __restrict__ phi, float* __restrict__ psi,
                                                         float* __restrict__ grad, int32_t* __restrict__ label,
                                                         int32_t* __restrict__ second) {
    const int k = blockIdx.x * kBlock + threadIdx.x;
    if (k >= m) {
        return;
    }
    const P3 p = p3(X[3 * k], X[3 * k + 1], X[3 * k + 2]);
    if (phi) {
        phi[k] = phi_point(f, a, p);
    }
    if (psi || grad) {
        const float4 g = psi_grad_point(f.self, a, b, p);
        if (psi) {
            psi[k] = g.w;
        }
        if (grad) {
            grad[3 * k] = g.x;
            grad[3 * k + 1] = g.y;
            grad[3 * k + 2] = g.z;
        }
    }
    if (label || second) {
        const int2 o = owner_at(f, p);
        if (label) {
            label[k] = o.x;
        }
        if (second) {
            second[k] = o.y;
        }
    }
}

// ---- control (one block) ---------------------------------------------------------------

struct CtlArgs {
    int n, max_iters, fire, rebuild_div;
    float dptol;
    int nparts;
    int force_serial;   // take the serial power sum every FIRE iteration (the exact path, for its tests)
};

// FIRE's power P = sum pw, whose sign (P >= 0) drives FireCtl::update
// (src/v2m_particles.h:58-73). The CPU path sums pw serially in double
// (src/v2m_particles.cpp:849-855). Here the partial sums are combined in a fixed
// tree; with A = sum |pw|, both that sum and the serial one are within
// (n + 64) * 2^-53 * A of the exact sum, so when |S| exceeds twice that bound both
// have the exact sum's sign. Otherwise the serial sum is recomputed here, in index
// order, exactly as the CPU does.
__device__ __forceinline__ void k_control_body(CtlArgs ca, Ctl* __restrict__ ctl, const double2* __restrict__ partials, const float* __restrict__ pw, const int32_t* __restrict__ slot_of, int bid, int nblk) {
    __shared__ double red[kBlock][2];
    __shared__ float chunk[kScanTile];
    __shared__ int use_serial;
    const int tid = threadIdx.x;
    double s = 0.0, a = 0.0;
    if (ca.fire) {
        for (int k = tid; k < ca.nparts; k += kBlock) {
            const double2 p = partials[k];
            s += p.x;
            a += p.y;
        }
    }
    red[tid][0] = s;
    red[tid][1] = a;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if (tid < w) {
            red[tid][0] += red[tid + w][0];
            red[tid][1] += red[tid + w][1];
        }
        __syncthreads();
    }
    double P = red[0][0];
    if (tid == 0) {
        const double A = red[0][1];
        const double bound = 2.0 * ((double)ca.n + 64.0) * 0x1p-53 * A;
        use_serial = ca.fire && (ca.force_serial || (A > 0.0 && !(fabs(P) > bound)));
    }
    __syncthreads();
    if (use_serial) {   // the CPU's own summation, in node index (id) order
        double pt = 0.0;
        for (int c0 = 0; c0 < ca.n; c0 += kScanTile) {
            for (int k = tid; k < kScanTile; k += kBlock) {
                chunk[k] = c0 + k < ca.n ? pw[slot_of[c0 + k]] : 0.0f;
            }
            __syncthreads();
            if (tid == 0) {
                const int m = ca.n - c0 < kScanTile ? ca.n - c0 : kScanTile;
                for (int k = 0; k < m; ++k) {
                    pt += chunk[k];
                }
            }
            __syncthreads();
        }
        P = pt;
    }
    if (tid != 0) {
        return;
    }
    const int it = ctl->it;
    if (use_serial) {
        ctl->power_fallbacks += 1;
    }
    ctl->power = P;
    if (ca.fire) {   // FireCtl::update
        if (P >= 0.0) {
            ctl->zero_v = 0;
            if (++ctl->npos > 5) {
                const float up = ctl->dt * 1.1f;
                ctl->dt = up < ctl->dtmax ? up : ctl->dtmax;
                ctl->alpha *= 0.99f;
            }
        } else {
            ctl->zero_v = 1;
            ctl->npos = 0;
            const float down = ctl->dt * 0.5f, lo = 0.05f * ctl->dt0;
            ctl->dt = down < lo ? lo : down;
            ctl->alpha = 0.1f;
            ctl->fire_resets += 1;
        }
    }
    ctl->it = it + 1;
    ctl->last_move = __uint_as_float(ctl->mmax_bits);
    ctl->fire_dt = ca.fire ? ctl->dt : 0.0f;
    int acc = 0, b99 = 31;
    for (int b = 0; b < 32; ++b) {
        acc += (int)ctl->hist[b];
        if ((double)acc >= 0.99 * (double)ca.n) {
            b99 = b;
            break;
        }
    }
    const float p99 = ldexpf(1.0f, b99 - 20 + 1);
    ctl->last_p99 = p99;
    const int done = p99 < ca.dptol && (!ca.fire || ctl->dt >= 0.25f * ctl->dt0);
    ctl->done = done;
    ctl->rebuild = 0;
    if (!done) {
        const int endgame = it >= ca.max_iters * 4 / 5 || p99 < 4.0f * ca.dptol;
        const int div = endgame ? (ca.rebuild_div > 1000 ? ca.rebuild_div : 1000) : ca.rebuild_div;
        ctl->rebuild = ctl->nhalf > ca.n / div;
    }
    for (int b = 0; b < 32; ++b) {
        ctl->hist[b] = 0;
    }
    ctl->mmax_bits = 0;
}

__global__ __launch_bounds__(kBlock) void k_control(CtlArgs ca, Ctl* __restrict__ ctl, const double2* __restrict__ partials, const float* __restrict__ pw, const int32_t* __restrict__ slot_of) {
    k_control_body(ca, ctl, partials, pw, slot_of, blockIdx.x, gridDim.x);
}

// clears the per-iteration counters that k_stats and k_refresh_list fill
__device__ __forceinline__ void k_iter_begin_body(Ctl* __restrict__ ctl, int bid, int nblk) {
    if (threadIdx.x == 0) {
        ctl->nhalf = 0;
        ctl->nrefresh = 0;
        ctl->nmove = 0;
        ctl->nmove_int = 0;
        ctl->nmove_rest = 0;
        ctl->nmove_def = 0;
    }
}

__global__ void k_iter_begin(Ctl* __restrict__ ctl) {
    k_iter_begin_body(ctl, blockIdx.x, gridDim.x);
}

// ---- grouped launches: one launch runs an iteration's kernel for many plans -------
//
// Each plan's arguments sit in a device table; a block finds its plan and its block
// index within that plan's grid (bplan / bfirst, per grid shape) and runs the same
// body a single-plan launch of that grid runs, so every plan computes exactly what
// its own launches would.

enum GridShape { kShNodes, kShLean, kShGen, kShParts, kShRefresh, kShOne, kShapes };

struct PlanIter {
    Field f;
    Hash H;
    TrussIn tin;
    MoveArgs ma;
    CtlArgs ca;
    int n, fire, nparts, voxel;   // voxel: the voxel-trap move (no queues)
    float fscale, fsurf, skin;
    Ctl* ctl;
    int32_t* nbrT;
    int32_t* nnb;
    float4 *X, *V, *F, *P0;
    uint8_t* typ;
    uint32_t* part;
    const uint16_t* lab;
    float *mv, *pw;
    const float* thr;
    uint8_t* flag;
    double2* partials;
    int32_t *rqueue, *mq, *mqi, *mq2, *mqd;
    const int32_t* slot_of;
    uint8_t* sticky;   // handed back to the general pass before (k_force queues them there)
    int nblk[kShapes];
};

#define TN_PLAN_BLOCK(shape)                              \
    const int v_ = bplan[blockIdx.x];                     \
    const PlanIter& P = tab[v_];                          \
    const int bid_ = (int)blockIdx.x - bfirst[v_];        \
    const int nb_ = P.nblk[shape]

__global__ void kg_iter_begin(const PlanIter* __restrict__ tab, const int32_t* __restrict__ bplan,
                              const int32_t* __restrict__ bfirst) {
    TN_PLAN_BLOCK(kShOne);
    k_iter_begin_body(P.ctl, bid_, nb_);
}

__global__ __launch_bounds__(kBlock) void kg_force(const PlanIter* __restrict__ tab, const int32_t* __restrict__ bplan,
                                                   const int32_t* __restrict__ bfirst) {
    TN_PLAN_BLOCK(kShNodes);
    k_force_body(P.X, P.typ, P.nbrT, P.nnb, P.fscale, P.fsurf, P.n, P.F, P.ctl, P.voxel ? nullptr : P.mq2, P.flag,
                 P.sticky, bid_, nb_);
}

template <bool kPlain>
__global__ __launch_bounds__(kBlock) void kg_move_easy(const PlanIter* __restrict__ tab,
                                                       const int32_t* __restrict__ bplan,
                                                       const int32_t* __restrict__ bfirst) {
    TN_PLAN_BLOCK(kShNodes);
    if (P.voxel) {   // the voxel-trap move's pass 1; pass 2 takes the interface queue's place
        k_move_voxel_body(P.f, P.ma, P.ctl, P.n, P.F, P.X, P.V, P.typ, P.part, P.lab, P.mv, P.pw, P.mq, &P.ctl->nmove,
                          bid_);
        return;
    }
    k_move_easy_body<kPlain>(P.f, P.ma, P.ctl, P.n, P.F, P.X, P.V, P.typ, P.part, P.lab, P.mv, P.pw, P.mq, P.mqi, P.flag,
                             bid_, nb_);
}

// the four move passes: kQueue 0 = the general queue, 1 = interior (lean), 2 =
// interface (lean), 3 = the handed-back nodes (general)
template <int kQueue, bool kPlain>
__global__ __launch_bounds__(kBlock, TN_MOVE_WAVES) void kg_move(const PlanIter* __restrict__ tab,
                                                                 const int32_t* __restrict__ bplan,
                                                                 const int32_t* __restrict__ bfirst) {
    constexpr int kMode = kQueue == 1 ? kLeanInterior : (kQueue == 2 ? kLeanInterface : kGeneral);
    TN_PLAN_BLOCK(kMode == kGeneral ? kShGen : kShLean);
    if (P.voxel) {   // the voxel-trap move's pass 2 (its only queue)
        if (kQueue == 2) {
            k_move_voxel_wave_body(P.f, P.ma, P.ctl, P.mq, &P.ctl->nmove, P.F, P.X, P.V, P.typ, P.part, P.lab, P.mv,
                                   P.pw, bid_, nb_);
        }
        return;
    }
    const int32_t* queue = kQueue == 0 ? P.mq2 : (kQueue == 1 ? P.mqi : (kQueue == 2 ? P.mq : P.mqd));
    const int32_t* count = kQueue == 0 ? &P.ctl->nmove_rest
                                       : (kQueue == 1 ? &P.ctl->nmove_int
                                                      : (kQueue == 2 ? &P.ctl->nmove : &P.ctl->nmove_def));
    const bool lean = kMode != kGeneral;
    k_move_body<kMode, kPlain>(P.f, P.ma, P.ctl, queue, P.F, P.X, P.V, P.typ, P.part, P.lab, P.mv, P.pw, count,
                       lean ? &P.ctl->nmove_def : nullptr, lean ? P.mqd : nullptr, lean ? P.sticky : nullptr, bid_,
                       nb_);
}

__global__ __launch_bounds__(kBlock) void kg_stats(const PlanIter* __restrict__ tab, const int32_t* __restrict__ bplan,
                                                   const int32_t* __restrict__ bfirst) {
    TN_PLAN_BLOCK(kShParts);
    k_stats_body(P.X, P.P0, P.mv, P.pw, P.n, P.skin, P.fire, P.thr, P.ctl, P.flag, P.partials, bid_, nb_);
}

__global__ __launch_bounds__(kBlock) void kg_refresh_list(const PlanIter* __restrict__ tab,
                                                          const int32_t* __restrict__ bplan,
                                                          const int32_t* __restrict__ bfirst) {
    TN_PLAN_BLOCK(kShNodes);
    k_refresh_list_body(P.flag, P.n, P.ctl, P.rqueue, bid_, nb_);
}

__global__ __launch_bounds__(kWave) void kg_truss_node(const PlanIter* __restrict__ tab,
                                                       const int32_t* __restrict__ bplan,
                                                       const int32_t* __restrict__ bfirst) {
    TN_PLAN_BLOCK(kShRefresh);
    k_truss_node_body(P.H, P.tin, P.n, P.rqueue, P.ctl, P.nbrT, P.nnb, P.P0, bid_, nb_);
}

__global__ __launch_bounds__(kBlock) void kg_control(const PlanIter* __restrict__ tab,
                                                     const int32_t* __restrict__ bplan,
                                                     const int32_t* __restrict__ bfirst) {
    TN_PLAN_BLOCK(kShOne);
    k_control_body(P.ca, P.ctl, P.partials, P.pw, P.slot_of, bid_, nb_);
}

// every plan's control block into one array (one copy to the host)
__global__ void kg_ctl_gather(const PlanIter* __restrict__ tab, const int32_t* __restrict__ bplan,
                              const int32_t* __restrict__ bfirst, Ctl* __restrict__ out) {
    TN_PLAN_BLOCK(kShOne);
    (void)bid_;
    (void)nb_;
    const uint32_t* src = reinterpret_cast<const uint32_t*>(P.ctl);
    uint32_t* dst = reinterpret_cast<uint32_t*>(out + v_);
    static_assert(sizeof(Ctl) % 4 == 0, "Ctl is copied in words");
    for (int k = threadIdx.x; k < (int)(sizeof(Ctl) / 4); k += blockDim.x) {
        dst[k] = src[k];
    }
}

#undef TN_PLAN_BLOCK

__global__ void k_order_begin(Ctl* __restrict__ ctl) {
    if (threadIdx.x < 4) {
        ctl->order_cnt[threadIdx.x] = 0;
    }
}

}  // namespace tnk
