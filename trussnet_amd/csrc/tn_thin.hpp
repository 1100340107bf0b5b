// tn_thin.hpp -- node thinning and feature protection on the device: thin_nodes
// (src/v2m_particles.cpp:1223-1357) and the per-node part of protect_features
// (src/v2m_particles.cpp:1002-1221).
//
// thin_nodes is a greedy pass: junction and corner nodes are kept; then the
// interface and interior nodes, interfaces first, by h, by index, are each kept
// unless a node kept before it lies within thin * h of it on the same interface
// (or, for an interior node, shares its label), searched in the 27 cells of the
// node's level of a multi-level hash. Here the greedy order becomes a rank per
// node, each node's list of earlier conflicting nodes is built once (a superset
// grid per level, then the reference's exact cell window, distance and label
// tests), and rounds decide every node whose earlier conflicts are all decided:
// removed when one of them is kept, kept when all are removed. Decisions never
// change once made, so a round may read states the same round wrote; each round
// decides at least the first undecided node in the order, and the result is the
// sequential pass's.
#pragma once

#include "tn_kernels.hpp"

namespace tnthin {

using tnd::Field;
using tnk::kBlock;

constexpr int32_t kUndecided = 0, kKept = 1, kRemoved = 2;

// the nodes of a thinning pass, by index
struct Nodes {
    const float4* __restrict__ X;        // position and h at the node (v2m_h_at)
    const uint16_t* __restrict__ lab;
    const uint8_t* __restrict__ typ;
    const uint16_t* __restrict__ part;   // [2n]
    const uint16_t* __restrict__ part3;
    const int32_t* __restrict__ rank;    // position in the greedy order; -1: junction / corner (kept first)
    int n;
};

// thin_nodes's hash (src/v2m_particles.cpp:1240-1255) and the superset grids of its levels
struct Levels {
    float B, c0;
    int nlev;
    float lo[3];                       // the node coordinates' lower bound
    float C[12];                       // superset cell size per level
    int dim[12][3];                    // superset grid dims per level
    int64_t off[12];                   // the level's first cell in the cell arrays
};

__device__ __forceinline__ int level_of(const Levels& Lv, float r) {
    int L = 0;
    while (L + 1 < Lv.nlev && Lv.c0 * (float)(1 << L) < r) {
        ++L;
    }
    return L;
}

// the coordinate of the reference's cell along one axis, as its key holds it (20 bits)
__device__ __forceinline__ uint32_t ref_coord(float x, float c) {
    return (uint32_t)((int64_t)floorf(x / c) + (1 << 19)) & 0xfffffu;
}

// the two coordinates are in neighbouring (or the same) reference cells
__device__ __forceinline__ bool ref_adjacent(uint32_t a, uint32_t b) {
    const uint32_t d = (a - b) & 0xfffffu;
    return d == 0u || d == 1u || d == 0xfffffu;
}

__device__ __forceinline__ int grid_coord(const Levels& Lv, int L, int ax, float x) {
    const int c = (int)floorf((x - Lv.lo[ax]) / Lv.C[L]);
    const int m = Lv.dim[L][ax];
    return c < 0 ? 0 : (c >= m ? m - 1 : c);
}

// label l among node j's labels (the `has` of src/v2m_particles.cpp:1267-1270)
__device__ __forceinline__ bool has_label(const Nodes& nd, int j, int l) {
    const int t = nd.typ[j];
    return nd.lab[j] == l || (t != TN_INTERIOR && nd.part[2 * j] == l) || (t >= TN_JUNCTION && nd.part[2 * j + 1] == l) ||
           (t == TN_CORNER && nd.part3[j] == l);
}

// X (position, h at the node), the candidates' flags and sort keys, the start states
__global__ __launch_bounds__(kBlock) void k_thin_init(Field f, const float* __restrict__ P, const uint8_t* __restrict__ typ,
                                                      int n, float4* __restrict__ X, int32_t* __restrict__ cand,
                                                      int32_t* __restrict__ hkey, int32_t* __restrict__ state) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) {
        const float x = P[3 * i], y = P[3 * i + 1], z = P[3 * i + 2];
        const float h = tnd::size_at(f, x, y, z);
        X[i] = make_float4(x, y, z, h);
        const int t = typ[i];
        cand[i] = t == TN_INTERFACE || t == TN_INTERIOR ? 1 : 0;
        hkey[i] = (int32_t)__float_as_uint(h);   // h >= 0: its bits order as the floats
        state[i] = t >= TN_JUNCTION ? kKept : kUndecided;
    }
}

// keys[t] = key of the candidate at t: h (pass 1), then interior-ness (pass 2)
__global__ __launch_bounds__(kBlock) void k_thin_gather_keys(const int32_t* __restrict__ src, const int32_t* __restrict__ vals,
                                                             int m, int32_t* __restrict__ keys) {
    const int t = blockIdx.x * kBlock + threadIdx.x;
    if (t < m) {
        keys[t] = src[vals[t]];
    }
}

__global__ __launch_bounds__(kBlock) void k_thin_interior(const uint8_t* __restrict__ typ, int n, int32_t* __restrict__ ti) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) {
        ti[i] = typ[i] == TN_INTERIOR ? 1 : 0;
    }
}

// rank[vals[t]] = t; non-candidates keep -1
__global__ __launch_bounds__(kBlock) void k_thin_rank(const int32_t* __restrict__ vals, int m, int32_t* __restrict__ rank) {
    const int t = blockIdx.x * kBlock + threadIdx.x;
    if (t < m) {
        rank[vals[t]] = t;
    }
}

// every node into the superset grid of each level: cell counts
__global__ __launch_bounds__(kBlock) void k_thin_count_cells(Levels Lv, const float4* __restrict__ X, int n, int L,
                                                             int32_t* __restrict__ cnt) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) {
        const float4 x = X[i];
        const int64_t c = Lv.off[L] + grid_coord(Lv, L, 0, x.x) +
                          (int64_t)Lv.dim[L][0] * (grid_coord(Lv, L, 1, x.y) + (int64_t)Lv.dim[L][1] * grid_coord(Lv, L, 2, x.z));
        atomicAdd(&cnt[c], 1);
    }
}

__global__ __launch_bounds__(kBlock) void k_thin_fill_cells(Levels Lv, const float4* __restrict__ X, int n, int L,
                                                            const int32_t* __restrict__ start, int32_t* __restrict__ fill,
                                                            int32_t* __restrict__ list) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) {
        const float4 x = X[i];
        const int64_t c = Lv.off[L] + grid_coord(Lv, L, 0, x.x) +
                          (int64_t)Lv.dim[L][0] * (grid_coord(Lv, L, 1, x.y) + (int64_t)Lv.dim[L][1] * grid_coord(Lv, L, 2, x.z));
        list[start[c] + atomicAdd(&fill[c], 1)] = i;
    }
}

// Walk the nodes that can crowd candidate i (kept before it if kept at all): within
// its reference cell window, closer than thin * h_i, on its interface or with its
// label. body(j) is called for each; it returns false to stop.
template <class Body>
__device__ __forceinline__ void thin_conflicts(const Levels& Lv, const Nodes& nd, const int32_t* __restrict__ cstart,
                                               const int32_t* __restrict__ clist, int i, Body body) {
    const float4 xi = nd.X[i];
    const float r = Lv.B * xi.w, r2 = r * r;
    const int L = level_of(Lv, r);
    const float c = Lv.c0 * (float)(1 << L);
    const uint32_t rx = ref_coord(xi.x, c), ry = ref_coord(xi.y, c), rz = ref_coord(xi.z, c);
    const int ri = nd.rank[i];
    const bool iface = nd.typ[i] == TN_INTERFACE;
    const int a = nd.lab[i], b = iface ? (int)nd.part[2 * i] : -1;
    // the reference's window is the 27 cells of size c around i's: within 2c along each
    // axis (2.5c: room for the rounding of the cell coordinates)
    const float w = 2.5f * c;
    const int x0 = grid_coord(Lv, L, 0, xi.x - w), x1 = grid_coord(Lv, L, 0, xi.x + w);
    const int y0 = grid_coord(Lv, L, 1, xi.y - w), y1 = grid_coord(Lv, L, 1, xi.y + w);
    const int z0 = grid_coord(Lv, L, 2, xi.z - w), z1 = grid_coord(Lv, L, 2, xi.z + w);
    for (int gz = z0; gz <= z1; ++gz) {
        for (int gy = y0; gy <= y1; ++gy) {
            const int64_t row = Lv.off[L] + (int64_t)Lv.dim[L][0] * (gy + (int64_t)Lv.dim[L][1] * gz);
            const int s0 = cstart[row + x0], s1 = cstart[row + x1 + 1];
            for (int s = s0; s < s1; ++s) {
                const int j = clist[s];
                const int rj = nd.rank[j];
                if (j == i || !(rj < ri)) {   // kept before i: junction / corner (-1) or earlier in the order
                    continue;
                }
                const float4 xj = nd.X[j];
                if (!ref_adjacent(ref_coord(xj.x, c), rx) || !ref_adjacent(ref_coord(xj.y, c), ry) ||
                    !ref_adjacent(ref_coord(xj.z, c), rz)) {
                    continue;
                }
                const float ex = xj.x - xi.x, ey = xj.y - xi.y, ez = xj.z - xi.z;
                if (ex * ex + ey * ey + ez * ez >= r2) {
                    continue;
                }
                if (iface ? (nd.typ[j] != TN_INTERIOR && has_label(nd, j, a) && has_label(nd, j, b)) : has_label(nd, j, a)) {
                    if (!body(j)) {
                        return;
                    }
                }
            }
        }
    }
}

__global__ __launch_bounds__(kBlock) void k_thin_count(Levels Lv, Nodes nd, const int32_t* __restrict__ cstart,
                                                       const int32_t* __restrict__ clist, const int32_t* __restrict__ cands,
                                                       int m, int32_t* __restrict__ ncon) {
    const int t = blockIdx.x * kBlock + threadIdx.x;
    if (t < m) {
        int k = 0;
        thin_conflicts(Lv, nd, cstart, clist, cands[t], [&](int) {
            ++k;
            return true;
        });
        ncon[t] = k;
    }
}

__global__ __launch_bounds__(kBlock) void k_thin_fill(Levels Lv, Nodes nd, const int32_t* __restrict__ cstart,
                                                      const int32_t* __restrict__ clist, const int32_t* __restrict__ cands,
                                                      int m, const int32_t* __restrict__ coff, int32_t* __restrict__ con) {
    const int t = blockIdx.x * kBlock + threadIdx.x;
    if (t < m) {
        int k = coff[t];
        thin_conflicts(Lv, nd, cstart, clist, cands[t], [&](int j) {
            con[k++] = j;
            return true;
        });
    }
}

// One round over the undecided candidates (cands[t], their conflicts at con[coff[t]..coff[t+1])).
// *left counts the ones still undecided after it.
__global__ __launch_bounds__(kBlock) void k_thin_round(const int32_t* __restrict__ cands, int m,
                                                       const int32_t* __restrict__ coff, const int32_t* __restrict__ con,
                                                       int32_t* state, int32_t* __restrict__ left) {
    const int t = blockIdx.x * kBlock + threadIdx.x;
    bool open = false;
    if (t < m) {
        const int i = cands[t];
        if (__atomic_load_n(&state[i], __ATOMIC_RELAXED) == kUndecided) {
            bool wait = false, crowded = false;
            for (int e = coff[t]; e < coff[t + 1]; ++e) {
                const int32_t s = __atomic_load_n(&state[con[e]], __ATOMIC_RELAXED);
                if (s == kKept) {
                    crowded = true;
                    break;
                }
                wait = wait || s == kUndecided;
            }
            if (crowded) {
                __atomic_store_n(&state[i], kRemoved, __ATOMIC_RELAXED);
            } else if (!wait) {
                __atomic_store_n(&state[i], kKept, __ATOMIC_RELAXED);
            } else {
                open = true;
            }
        }
    }
    const uint64_t mask = __ballot(open);
    if ((threadIdx.x & (tnk::kWave - 1)) == 0 && mask) {
        atomicAdd(left, __popcll(mask));
    }
}

// keep flags (as counts for the scan) from the states
__global__ __launch_bounds__(kBlock) void k_thin_keep(const int32_t* __restrict__ state, int n, int32_t* __restrict__ keep) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) {
        keep[i] = state[i] == kKept ? 1 : 0;
    }
}

// protect_features's test of each node (src/v2m_particles.cpp:1166-1190): a node other
// than a corner is dropped when it lies inside one of the balls listed for its cell
// of size cs. cell_key[0..ncell) is sorted; cell_ball the ball of each entry.
__global__ __launch_bounds__(kBlock) void k_protect(const float* __restrict__ P, const uint8_t* __restrict__ typ, int n,
                                                    float cs, const uint64_t* __restrict__ cell_key,
                                                    const int32_t* __restrict__ cell_ball, int ncell,
                                                    const float4* __restrict__ balls, int32_t* __restrict__ keep) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) {
        return;
    }
    int k = 1;
    if (typ[i] != TN_CORNER) {
        const float px = P[3 * i], py = P[3 * i + 1], pz = P[3 * i + 2];
        const uint64_t key = (uint64_t)((int64_t)floorf(px / cs) + (1 << 20)) << 42 |
                             (uint64_t)((int64_t)floorf(py / cs) + (1 << 20)) << 21 |
                             (uint64_t)((int64_t)floorf(pz / cs) + (1 << 20));
        int lo = 0, hi = ncell;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (cell_key[mid] < key) {
                lo = mid + 1;
            } else {
                hi = mid;
            }
        }
        for (int e = lo; e < ncell && cell_key[e] == key; ++e) {
            const float4 bl = balls[cell_ball[e]];
            const float dx = px - bl.x, dy = py - bl.y, dz = pz - bl.z;
            if (dx * dx + dy * dy + dz * dz < bl.w) {
                k = 0;
                break;
            }
        }
    }
    keep[i] = k;
}

// the kept nodes, compacted in index order (offs: the exclusive scan of the keep flags)
__global__ __launch_bounds__(kBlock) void k_compact_nodes(const int32_t* __restrict__ offs, int n, const float* __restrict__ P,
                                                          const uint16_t* __restrict__ lab, const uint8_t* __restrict__ typ,
                                                          const uint16_t* __restrict__ part, const uint16_t* __restrict__ part3,
                                                          float* __restrict__ oP, uint16_t* __restrict__ olab,
                                                          uint8_t* __restrict__ otyp, uint16_t* __restrict__ opart,
                                                          uint16_t* __restrict__ opart3) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n && offs[i + 1] != offs[i]) {
        const int w = offs[i];
        oP[3 * w] = P[3 * i];
        oP[3 * w + 1] = P[3 * i + 1];
        oP[3 * w + 2] = P[3 * i + 2];
        olab[w] = lab[i];
        otyp[w] = typ[i];
        opart[2 * w] = part[2 * i];
        opart[2 * w + 1] = part[2 * i + 1];
        opart3[w] = part3[i];
    }
}

}  // namespace tnthin
