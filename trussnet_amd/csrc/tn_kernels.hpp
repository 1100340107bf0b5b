// tn_kernels.hpp -- the relaxation's gfx950 kernels.
//
// Rebuild (src/v2m_particles.cpp:771-820): k_keys (cell key + count per cell),
// k_scan_* (exclusive scan: cell starts), k_radix_* (stable LSD radix sort of
// the node ids by cell key: ascending id within a cell, as the CPU's serial
// counting sort), k_gather (positions in cell order), k_neighbors (K-nearest
// lists), k_order_* (move order grouped by node type).
// Iteration (src/v2m_particles.cpp:828-947): k_force, k_move, k_stats (histogram,
// max, Verlet count, FIRE power partial sums), k_refresh (own-list refresh of
// nodes past a full skin), k_control (one block: FIRE's update, p99, the stop
// and rebuild decisions, all on the device).
#pragma once

#include "tn_device.hpp"

namespace tnk {

using namespace tnd;

constexpr int kWave = 64;
constexpr int kBlock = 256;
constexpr int kNbrBlock = 128;   // K-nearest search: 128 lanes x 24 x 8 B of LDS K-lists = 24 KB

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
    // FIRE (FireCtl, src/v2m_particles.h:47-76)
    float dt0, dt, dtmax, alpha;
    int32_t npos;
    // accumulators, cleared by k_control
    uint32_t hist[32];
    uint32_t mmax_bits;
    int32_t order_cnt[4];
    int32_t order_cur[4];
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

__global__ __launch_bounds__(kBlock) void k_keys(Field f, Hash H, const float4* __restrict__ X, int n, float tskin,
                                                 int32_t* __restrict__ key, int32_t* __restrict__ cnt) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) {
        return;
    }
    const float4 x = X[i];
    const float h = size_at(f, x.x, x.y, x.z);
    const int k = cell_key(H, level_for(H, tskin * h), x.x, x.y, x.z);
    key[i] = k;
    atomicAdd(&cnt[k], 1);
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

// ---- gather (positions in cell order) + P0 = P ---------------------------------------

__global__ __launch_bounds__(kBlock) void k_gather(const int32_t* __restrict__ sorted, const float4* __restrict__ X,
                                                   int n, float4* __restrict__ Ps, float4* __restrict__ P0) {
    const int s = blockIdx.x * kBlock + threadIdx.x;
    if (s < n) {
        Ps[s] = X[sorted[s]];
        P0[s] = X[s];
    }
}

// ---- K-nearest lists ----------------------------------------------------------------

// nodes visited in cell order: a wave's lanes scan the same cells
__global__ __launch_bounds__(kNbrBlock) void k_neighbors(Hash H, TrussIn in, int n, int32_t* __restrict__ nbrT,
                                                         int32_t* __restrict__ nnb) {
    __shared__ float ds[kK * kNbrBlock];
    __shared__ int ids[kK * kNbrBlock];
    const int s = blockIdx.x * kNbrBlock + threadIdx.x;
    if (s >= n) {
        return;
    }
    truss_of(H, in, in.sorted[s], n, ds + threadIdx.x, ids + threadIdx.x, kNbrBlock, nbrT, nnb);
}

// own-list refresh of the nodes queued by k_stats (v2m_particles.cpp:925-931):
// stale cells and Ps, current positions and types; then P0 = P for them
__global__ __launch_bounds__(kNbrBlock) void k_refresh(Hash H, TrussIn in, int n, const int32_t* __restrict__ queue,
                                                       const Ctl* __restrict__ ctl, int32_t* __restrict__ nbrT,
                                                       int32_t* __restrict__ nnb, float4* __restrict__ P0) {
    __shared__ float ds[kK * kNbrBlock];
    __shared__ int ids[kK * kNbrBlock];
    const int m = ctl->nrefresh;
    for (int q = blockIdx.x * kNbrBlock + threadIdx.x; q < m; q += gridDim.x * kNbrBlock) {
        const int i = queue[q];
        truss_of(H, in, i, n, ds + threadIdx.x, ids + threadIdx.x, kNbrBlock, nbrT, nnb);
        P0[i] = in.X[i];
    }
}

// ---- move order grouped by type: interface, junction, interior, corner ---------------

__device__ __forceinline__ int type_rank(int t) {
    return t == TN_INTERFACE ? 0 : (t == TN_JUNCTION ? 1 : (t == TN_INTERIOR ? 2 : 3));
}

__global__ __launch_bounds__(kBlock) void k_order_count(const uint8_t* __restrict__ typ, int n, Ctl* __restrict__ ctl) {
    __shared__ int c[4];
    if (threadIdx.x < 4) {
        c[threadIdx.x] = 0;
    }
    __syncthreads();
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) {
        atomicAdd(&c[type_rank(typ[i] & 3)], 1);
    }
    __syncthreads();
    if (threadIdx.x < 4 && c[threadIdx.x]) {
        atomicAdd(&ctl->order_cnt[threadIdx.x], c[threadIdx.x]);
    }
}

// (the order within a type group is not deterministic; it only steers performance:
// the move of node i touches node i alone)
__global__ __launch_bounds__(kBlock) void k_order_scatter(const uint8_t* __restrict__ typ, int n,
                                                          Ctl* __restrict__ ctl, int32_t* __restrict__ order) {
    __shared__ int c[4], b[4];
    if (threadIdx.x < 4) {
        c[threadIdx.x] = 0;
    }
    __syncthreads();
    const int i = blockIdx.x * kBlock + threadIdx.x;
    int r = 0, loc = 0;
    if (i < n) {
        r = type_rank(typ[i] & 3);
        loc = atomicAdd(&c[r], 1);
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        int start = 0;
        for (int k = 0; k < (int)threadIdx.x; ++k) {
            start += ctl->order_cnt[k];
        }
        b[threadIdx.x] = c[threadIdx.x] ? start + atomicAdd(&ctl->order_cur[threadIdx.x], c[threadIdx.x]) : 0;
    }
    __syncthreads();
    if (i < n) {
        order[b[r] + loc] = i;
    }
}

// ---- force --------------------------------------------------------------------------

__global__ __launch_bounds__(kBlock) void k_force(const float4* __restrict__ X, const uint8_t* __restrict__ typ,
                                                  const int32_t* __restrict__ nbrT, const int32_t* __restrict__ nnb,
                                                  float fscale, float fsurf, int n, float4* __restrict__ F) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) {
        F[i] = bar_force(X, typ, nbrT, nnb, fscale, fsurf, i, n);
    }
}

// ---- move (v2m_move, src/opencl/v2m_particle_body.cl:844-935) ----------------------

struct MoveArgs {
    float dt, maxstep, snap;
    int fire;
};

__device__ __forceinline__ uint32_t pack_part(int b, int c) {
    return (uint32_t)(b & 0xFFFF) | ((uint32_t)(c & 0xFFFF) << 16);
}

__global__ __launch_bounds__(kBlock) void k_move(Field f, MoveArgs ma, const Ctl* __restrict__ ctl,
                                                 const int32_t* __restrict__ order, int n,
                                                 const float4* __restrict__ F, float4* __restrict__ X,
                                                 float4* __restrict__ V, uint8_t* __restrict__ typ,
                                                 uint32_t* __restrict__ part, const uint16_t* __restrict__ lab,
                                                 float* __restrict__ mv, float* __restrict__ pw) {
    const int g = blockIdx.x * kBlock + threadIdx.x;
    if (g >= n) {
        return;
    }
    const int i = order[g];
    const float fdt = ctl->dt, falpha = ctl->alpha;
    const int fire = ma.fire;
    const float4 x = X[i];
    const float h = x.w;   // h at the node: v2m_h_at(P_i), kept current by every move
    const int ty0 = typ[i];
    float p_w = 0.0f;
    float4 vout = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    float m = 0.0f;
    if (ty0 != TN_CORNER) {
        const float4 Fi = F[i];
        const float stiff = Fi.w > 1.0f ? Fi.w : 1.0f;
        const float ac[3] = { Fi.x / stiff, Fi.y / stiff, Fi.z / stiff };
        float s[3];
        if (fire) {
            float v[3] = { 0.0f, 0.0f, 0.0f };
            if (!ctl->zero_v) {
                const float4 v4 = V[i];
                v[0] = v4.x;
                v[1] = v4.y;
                v[2] = v4.z;
            }
            const float pv = ac[0] * v[0] + ac[1] * v[1] + ac[2] * v[2];
            p_w = pv;
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
                s[k] = fdt * v[k];
            }
        } else {
            for (int k = 0; k < 3; ++k) {
                s[k] = ma.dt * ac[k];
            }
        }
        const float sl = sqrtf(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
        if (!(sl < 5e-4f * h)) {
            if (sl > ma.maxstep * h) {
                for (int k = 0; k < 3; ++k) {
                    s[k] *= ma.maxstep * h / sl;
                }
            }
            const uint32_t pp = part[i];
            NodeState ns;
            ns.p = p3(x.x, x.y, x.z);
            ns.hn = h;
            ns.ty = ty0;
            ns.b = (int)(pp & 0xFFFF);
            ns.c = (int)(pp >> 16);
            m = trapped_step(f, h, p3(s[0], s[1], s[2]), ma.maxstep, ma.snap, lab[i], ns);
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
            if (fire && m > 0.0f && ns.ty == ty0) {
                vout = make_float4((ns.p.x - x.x) / fdt, (ns.p.y - x.y) / fdt, (ns.p.z - x.z) / fdt, 0.0f);
            }
        }
    }
    if (fire) {
        V[i] = vout;
        pw[i] = p_w;
    }
    mv[i] = m;
}

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

__global__ __launch_bounds__(kBlock) void k_stats(const float4* __restrict__ X, const float4* __restrict__ P0,
                                                  const float* __restrict__ mv, const float* __restrict__ pw, int n,
                                                  float skin, int fire, const float* __restrict__ thr,
                                                  Ctl* __restrict__ ctl, int32_t* __restrict__ queue,
                                                  double2* __restrict__ partials) {
    __shared__ uint32_t lh[32];
    __shared__ uint32_t lmax;
    __shared__ int lhalf;
    __shared__ double lps[kBlock / kWave][2];
    const int tid = threadIdx.x;
    if (tid < 32) {
        lh[tid] = 0;
    }
    if (tid == 0) {
        lmax = 0;
        lhalf = 0;
    }
    __syncthreads();
    const int i = blockIdx.x * kBlock + tid;
    double ps = 0.0, pa = 0.0;
    if (i < n) {
        const float m = mv[i];
        atomicAdd(&lh[hist_bin(m, thr)], 1u);
        atomicMax(&lmax, __float_as_uint(m));
        const float4 x = X[i], x0 = P0[i];
        const float ex = x.x - x0.x, ey = x.y - x0.y, ez = x.z - x0.z;
        const float h = x.w;
        const float m2 = ex * ex + ey * ey + ez * ez, s2 = skin * skin * h * h;
        if (m2 > s2) {
            queue[atomicAdd(&ctl->nrefresh, 1)] = i;
        } else if (m2 > 0.25f * s2) {
            atomicAdd(&lhalf, 1);
        }
        if (fire) {
            ps = (double)pw[i];
            pa = fabs(ps);
        }
    }
    if (fire) {   // fixed-shape tree: the partial sums are reproducible
#pragma unroll
        for (int d = kWave / 2; d > 0; d >>= 1) {
            ps += __shfl_xor(ps, d, kWave);
            pa += __shfl_xor(pa, d, kWave);
        }
        if ((tid & (kWave - 1)) == 0) {
            lps[tid >> 6][0] = ps;
            lps[tid >> 6][1] = pa;
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
            const double s = (lps[0][0] + lps[1][0]) + (lps[2][0] + lps[3][0]);
            const double a = (lps[0][1] + lps[1][1]) + (lps[2][1] + lps[3][1]);
            partials[blockIdx.x] = make_double2(s, a);
        }
    }
}

// ---- control (one block) ---------------------------------------------------------------

struct CtlArgs {
    int n, max_iters, fire, rebuild_div;
    float dptol;
    int nparts;
};

// FIRE's power P = sum pw, whose sign (P >= 0) drives FireCtl::update
// (src/v2m_particles.h:58-73). The CPU path sums pw serially in double
// (src/v2m_particles.cpp:849-855). Here the partial sums are combined in a fixed
// tree; with A = sum |pw|, both that sum and the serial one are within
// (n + 64) * 2^-53 * A of the exact sum, so when |S| exceeds twice that bound both
// have the exact sum's sign. Otherwise the serial sum is recomputed here, in index
// order, exactly as the CPU does.
__global__ __launch_bounds__(kBlock) void k_control(CtlArgs ca, Ctl* __restrict__ ctl,
                                                    const double2* __restrict__ partials,
                                                    const float* __restrict__ pw) {
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
        use_serial = ca.fire && A > 0.0 && !(fabs(P) > bound);
    }
    __syncthreads();
    if (use_serial) {   // the CPU's own summation, in index order
        double pt = 0.0;
        for (int c0 = 0; c0 < ca.n; c0 += kScanTile) {
            for (int k = tid; k < kScanTile; k += kBlock) {
                chunk[k] = c0 + k < ca.n ? pw[c0 + k] : 0.0f;
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

// clears the per-iteration counters that k_stats and k_refresh consume
__global__ void k_iter_begin(Ctl* __restrict__ ctl) {
    if (threadIdx.x == 0) {
        ctl->nhalf = 0;
        ctl->nrefresh = 0;
    }
}

__global__ void k_order_begin(Ctl* __restrict__ ctl) {
    if (threadIdx.x < 4) {
        ctl->order_cnt[threadIdx.x] = 0;
        ctl->order_cur[threadIdx.x] = 0;
    }
}

}  // namespace tnk
