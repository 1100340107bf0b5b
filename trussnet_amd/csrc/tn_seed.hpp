// tn_seed.hpp -- graded HCP seeding on the device (the body of seed_cpu,
// src/v2m_particles.cpp:637-726: count -> scan -> fill over the voxels, then the
// classification of every seed).
//
// Each voxel claims the points of its seed level's global HCP lattice that fall
// in its box (v2m_seed_voxel, src/opencl/v2m_seed_body.cl:449-495), keeping
// those whose owning label is not the exterior; the fill writes them in voxel
// order at the scanned offsets, so the node order is the reference's. Each seed
// is then projected onto a nearby interface or junction curve
// (v2m_seed_classify, src/opencl/v2m_seed_body.cl:503-594).
#pragma once

#include "tn_device.hpp"

namespace tnk {

using namespace tnd;

// v2m_seed_level, src/opencl/v2m_seed_body.cl:441-443
__device__ __forceinline__ int seed_level(int grade, int nseed) {
    return nseed <= 1 ? 0 : (int)floorf((float)grade * (float)(nseed - 1) / 255.0f + 0.5f);
}

// v2m_hcp_point, src/opencl/v2m_seed_body.cl:427-432
__device__ __forceinline__ P3 hcp_point(float s, int i, int j, int k) {
    const float dy = s * 0.8660254f, dz = s * 0.8164966f;
    P3 p;
    p.x = i * s + ((j & 1) ? 0.5f * s : 0.0f) + ((k & 1) ? 0.5f * s : 0.0f);
    p.y = j * dy + ((k & 1) ? dy / 3.0f : 0.0f);
    p.z = k * dz;
    return p;
}

struct SeedArgs {
    const uint8_t* __restrict__ grade;
    const float* __restrict__ spacing;   // per seed level (v2m_seed_spacing, computed on the host)
    int nseed;
    int voxmode;   // the voxel trap: a point's label is its voxel's
};

// v2m_label_of with voxmode 1 (src/opencl/v2m_seed_body.cl:116-175): a brick with one
// label gives that label; otherwise the label of the voxel nearest the point
__device__ __forceinline__ int voxel_owner(const Field& f, P3 p) {
    const BrickLabels bl = brick_labels_at(f, p);
    if (bl.n == 1) {
        return brick_label(f, bl, 0);
    }
    return label_or_zero(f, (int)floorf(over_vx(f, p.x) + 0.5f), (int)floorf(over_vy(f, p.y) + 0.5f),
                         (int)floorf(over_vz(f, p.z) + 0.5f));
}

// v2m_seed_voxel, src/opencl/v2m_seed_body.cl:449-495. Write: store the kept
// points from slot `out` on. Returns the number kept.
template <bool kWrite>
__device__ int seed_voxel(const Field& f, const SeedArgs& sa, int i, int j, int k, float* __restrict__ P,
                          uint16_t* __restrict__ lab, int out) {
    const size_t v = vox(f, i, j, k);
    if (f.L[v] == 0) {
        return 0;
    }
    const float s = sa.spacing[seed_level(sa.grade[v], sa.nseed)];
    const float dy = s * 0.8660254f, dz = s * 0.8164966f;
    const float x0 = (i - 0.5f) * f.vx, x1 = (i + 0.5f) * f.vx;
    const float y0 = (j - 0.5f) * f.vy, y1 = (j + 0.5f) * f.vy;
    const float z0 = (k - 0.5f) * f.vz, z1 = (k + 0.5f) * f.vz;
    int n = 0;
    const int k0 = (int)ceilf(z0 / dz), k1 = (int)ceilf(z1 / dz);
    for (int kk = k0; kk < k1; ++kk) {
        const float oy = (kk & 1) ? dy / 3.0f : 0.0f;
        const int j0 = (int)ceilf((y0 - oy) / dy), j1 = (int)ceilf((y1 - oy) / dy);
        for (int jj = j0; jj < j1; ++jj) {
            const float ox = ((jj & 1) ? 0.5f * s : 0.0f) + ((kk & 1) ? 0.5f * s : 0.0f);
            const int i0 = (int)ceilf((x0 - ox) / s), i1 = (int)ceilf((x1 - ox) / s);
            for (int ii = i0; ii < i1; ++ii) {
                const P3 p = hcp_point(s, ii, jj, kk);
                const int l = sa.voxmode ? voxel_owner(f, p) : owner_at(f, p).x;
                if (l == 0 || l == kNoLab) {
                    continue;
                }
                if (kWrite) {
                    P[3 * (size_t)(out + n)] = p.x;
                    P[3 * (size_t)(out + n) + 1] = p.y;
                    P[3 * (size_t)(out + n) + 2] = p.z;
                    lab[out + n] = (uint16_t)l;
                }
                ++n;
            }
        }
    }
    return n;
}

__global__ __launch_bounds__(kBlock) void k_seed_count(Field f, SeedArgs sa, int32_t* __restrict__ cnt) {
    const int64_t v = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int64_t nv = (int64_t)f.nx * f.ny * f.nz;
    if (v >= nv) {
        return;
    }
    const int i = (int)(v % f.nx), j = (int)((v / f.nx) % f.ny), k = (int)(v / ((int64_t)f.nx * f.ny));
    cnt[v] = seed_voxel<false>(f, sa, i, j, k, nullptr, nullptr, 0);
}

__global__ __launch_bounds__(kBlock) void k_seed_fill(Field f, SeedArgs sa, const int32_t* __restrict__ offs,
                                                      float* __restrict__ P, uint16_t* __restrict__ lab) {
    const int64_t v = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int64_t nv = (int64_t)f.nx * f.ny * f.nz;
    if (v >= nv || offs[v + 1] == offs[v]) {
        return;
    }
    const int i = (int)(v % f.nx), j = (int)((v / f.nx) % f.ny), k = (int)(v / ((int64_t)f.nx * f.ny));
    seed_voxel<true>(f, sa, i, j, k, P, lab, offs[v]);
}

// v2m_seed_classify, src/opencl/v2m_seed_body.cl:503-594. (The size at the node is
// the nearest voxel's, clamped to the grid like v2m_h_at; a lattice point lies in
// its voxel's box, so the clamp never changes it.)
__global__ __launch_bounds__(kBlock) void k_seed_classify(Field f, int n, float* __restrict__ P,
                                                          const uint16_t* __restrict__ lab, uint8_t* __restrict__ typ,
                                                          uint16_t* __restrict__ part) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) {
        return;
    }
    const P3 p = p3(P[3 * (size_t)i], P[3 * (size_t)i + 1], P[3 * (size_t)i + 2]);
    typ[i] = TN_INTERIOR;
    part[2 * (size_t)i] = part[2 * (size_t)i + 1] = kNoLab;
    const int a = lab[i];
    const BrickLabels bl = brick_labels_at(f, p);
    if (bl.n < 2) {
        return;
    }
    int b = kNoLab;
    float pb = -1.0f;
    for (int s = 0; s < bl.n; ++s) {
        const int l = brick_label(f, bl, s);
        if (l != a) {
            const float v = phi_point(f, l, p);
            if (v > pb) {
                pb = v;
                b = l;
            }
        }
    }
    if (b == kNoLab || (b != 0 && b < a)) {
        return;
    }
    const float hs = size_at(f, p.x, p.y, p.z);
    const float4 g = psi_grad_point(f.self, a, b, p);
    const float gn = sqrtf(g.x * g.x + g.y * g.y + g.z * g.z);
    if (gn < 1e-9f || fabsf(g.w) / gn > 0.5f * hs) {
        return;
    }
    const P3ok pr = onto_interface(f.self, a, b, p, 0.6f * hs);
    if (!pr.ok) {
        return;
    }
    P3 q = pr.p;
    int ty = TN_INTERFACE;
    int c = kNoLab;
    float pc = -1.0f;
    const float pa = phi_point(f, a, q);
    for (int s = 0; s < bl.n; ++s) {
        const int l = brick_label(f, bl, s);
        if (l != a && l != b) {
            const float w = phi_point(f, l, q);
            if (w > pc) {
                pc = w;
                c = l;
            }
        }
    }
    int pc_lab = kNoLab;
    if (c != kNoLab && pc > 0.5f * pa) {
        const P3ok pj = onto_junction(f.self, a, b, c, q, 0.5f * hs);
        if (pj.ok) {
            q = pj.p;
            ty = TN_JUNCTION;
            pc_lab = c;
        }
    }
    typ[i] = (uint8_t)ty;
    part[2 * (size_t)i] = (uint16_t)b;
    part[2 * (size_t)i + 1] = (uint16_t)pc_lab;
    P[3 * (size_t)i] = q.x;
    P[3 * (size_t)i + 1] = q.y;
    P[3 * (size_t)i + 2] = q.z;
}

}  // namespace tnk

// ---- junction-line seeds and corner nodes (seed_junctions and the corner pass of
// seed_cpu_body, src/v2m_particles.cpp:80-244 and 686-726) --------------------------
//
// Both walk the grid vertices: vertex v in [0, (nx+1)(ny+1)(nz+1)) is (i, j, k) =
// (v % (nx+1) - 1, ...), the corner shared by voxels i..i+1, j..j+1, k..k+1. A
// cheap mark pass flags the vertices with 3 labels (junction candidates) or with 4+
// labels and no smaller 4+-label neighbour (corners); a scan turns the flags into
// a list in vertex order, and the costly projections then run on that dense list.

namespace tnk {

struct VertexGrid {
    int vx1, vy1;   // nx + 1, ny + 1
    int64_t nvert;
};

__device__ __forceinline__ void vertex_ijk(const VertexGrid& vg, int64_t v, int& i, int& j, int& k) {
    i = (int)(v % vg.vx1) - 1;
    j = (int)((v / vg.vx1) % vg.vy1) - 1;
    k = (int)(v / ((int64_t)vg.vx1 * vg.vy1)) - 1;
}

// v2m_vertex_nlab, src/opencl/v2m_seed_body.cl:613-630: the distinct labels of the
// 8 voxels at vertex (i,j,k) in order of first appearance; returns their count
__device__ __forceinline__ int vertex_labels(const Field& f, int i, int j, int k, int* labs) {
    int n = 0;
#pragma unroll
    for (int m = 0; m < 8; ++m) {
        const int l = label_or_zero(f, i + (m & 1), j + ((m >> 1) & 1), k + (m >> 2));
        bool seen = false;
        for (int x = 0; x < n; ++x) {
            seen |= labs[x] == l;
        }
        if (!seen) {
            labs[n++] = l;
        }
    }
    return n;
}

// the count alone, without a private array
__device__ __forceinline__ int vertex_label_count(const Field& f, int i, int j, int k) {
    int l[8];
#pragma unroll
    for (int m = 0; m < 8; ++m) {
        l[m] = label_or_zero(f, i + (m & 1), j + ((m >> 1) & 1), k + (m >> 2));
    }
    int n = 0;
#pragma unroll
    for (int m = 0; m < 8; ++m) {
        bool first = true;
#pragma unroll
        for (int x = 0; x < m; ++x) {
            first = first && l[x] != l[m];
        }
        n += first ? 1 : 0;
    }
    return n;
}

// v2m_corner_vertex's selection (src/opencl/v2m_seed_body.cl:632-651): 4+ labels
// and no 4+-label vertex among the 13 neighbours of smaller index
__device__ __forceinline__ bool corner_candidate(const Field& f, int i, int j, int k) {
    if (vertex_label_count(f, i, j, k) < 4) {
        return false;
    }
    for (int dz = -1; dz <= 0; ++dz) {
        for (int dy = -1; dy <= 1; ++dy) {
            for (int dx = -1; dx <= 1; ++dx) {
                const bool smaller = dz < 0 || (dy < 0 || (dy == 0 && dx < 0));
                if (smaller && vertex_label_count(f, i + dx, j + dy, k + dz) >= 4) {
                    return false;
                }
            }
        }
    }
    return true;
}

// kind 0: flag[v] = 3 labels at v (junction candidates); kind 1: corner vertices
template <int kKind>
__global__ __launch_bounds__(kBlock) void k_vertex_mark(Field f, VertexGrid vg, int32_t* __restrict__ flag) {
    const int64_t v = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (v >= vg.nvert) {
        return;
    }
    int i, j, k;
    vertex_ijk(vg, v, i, j, k);
    flag[v] = kKind == 0 ? (vertex_label_count(f, i, j, k) == 3 ? 1 : 0) : (corner_candidate(f, i, j, k) ? 1 : 0);
}

// list[offs[e]] = e for the flagged entries (offs: the exclusive scan of the flags)
__global__ __launch_bounds__(kBlock) void k_compact_index(const int32_t* __restrict__ offs, int64_t m,
                                                          int32_t* __restrict__ list) {
    const int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (e < m && offs[e + 1] != offs[e]) {
        list[offs[e]] = (int32_t)e;
    }
}

// A junction candidate: its position, labels (own, o0, o1) and the dedup key
// (seed level, cell x/y/z of cellfac * level spacing, sorted label triple), SoA.
struct JunctionCands {
    float* x;        // [3 m]
    int32_t* lab3;   // [3 m]
    int32_t* key;    // [7 m]: field e of candidate s at key[e * m + s]
    int32_t* ok;     // [m + 1]
    int m;
};

// v2m_junction_vertex, src/opencl/v2m_particle_body.cl:302-352, for one 3-label
// vertex of the list
__global__ __launch_bounds__(kBlock) void k_junction_cand(Field f, SeedArgs sa, float cellfac, VertexGrid vg,
                                                          const int32_t* __restrict__ list, JunctionCands jc) {
    const int s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= jc.m) {
        return;
    }
    jc.ok[s] = 0;
    int i, j, k;
    vertex_ijk(vg, list[s], i, j, k);
    int labs[8];
    vertex_labels(f, i, j, k, labs);
    P3 x = p3((i + 0.5f) * f.vx, (j + 0.5f) * f.vy, (k + 0.5f) * f.vz);
    int a = -1;
    float pa = -1.0f;
    for (int t = 0; t < 3; ++t) {   // own = the strongest non-zero label
        if (labs[t] != 0) {
            const float v = phi_point(f, labs[t], x);
            if (v > pa) {
                pa = v;
                a = labs[t];
            }
        }
    }
    int o[2], no = 0;
    for (int t = 0; t < 3; ++t) {
        if (labs[t] != a) {
            o[no++] = labs[t];
        }
    }
    const float vmin = fminf(fminf(f.vx, f.vy), f.vz);
    const P3ok pj = onto_junction(f.self, a, o[0], o[1], x, vmin);
    if (!pj.ok) {
        return;
    }
    x = pj.p;
    if (!labels_dominate(f, a, o[0], o[1], x)) {
        return;
    }
    const int vi = (int)floorf(over_vx(f, x.x) + 0.5f), vj = (int)floorf(over_vy(f, x.y) + 0.5f),
              vk = (int)floorf(over_vz(f, x.z) + 0.5f);
    if (vi < 0 || vj < 0 || vk < 0 || vi >= f.nx || vj >= f.ny || vk >= f.nz) {
        return;
    }
    const int lev = seed_level(sa.grade[vox(f, vi, vj, vk)], sa.nseed);
    const float c = cellfac * sa.spacing[lev];
    int t0 = a, t1 = o[0], t2 = o[1], tmp;   // the triple, ascending
    if (t1 < t0) { tmp = t0; t0 = t1; t1 = tmp; }
    if (t2 < t1) { tmp = t1; t1 = t2; t2 = tmp; }
    if (t1 < t0) { tmp = t0; t0 = t1; t1 = tmp; }
    const int m = jc.m;
    jc.key[s] = lev;
    jc.key[m + s] = (int)floorf(x.x / c);
    jc.key[2 * m + s] = (int)floorf(x.y / c);
    jc.key[3 * m + s] = (int)floorf(x.z / c);
    jc.key[4 * m + s] = t0;
    jc.key[5 * m + s] = t1;
    jc.key[6 * m + s] = t2;
    jc.x[3 * (size_t)s] = x.x;
    jc.x[3 * (size_t)s + 1] = x.y;
    jc.x[3 * (size_t)s + 2] = x.z;
    jc.lab3[3 * (size_t)s] = a;
    jc.lab3[3 * (size_t)s + 1] = o[0];
    jc.lab3[3 * (size_t)s + 2] = o[1];
    jc.ok[s] = 1;
}

// the valid candidates, in list order
__global__ __launch_bounds__(kBlock) void k_gather_cands(const JunctionCands jc, const int32_t* __restrict__ cl,
                                                         JunctionCands out) {
    const int s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= out.m) {
        return;
    }
    const int c = cl[s];
    for (int e = 0; e < 3; ++e) {
        out.x[3 * (size_t)s + e] = jc.x[3 * (size_t)c + e];
        out.lab3[3 * (size_t)s + e] = jc.lab3[3 * (size_t)c + e];
    }
    for (int e = 0; e < 7; ++e) {
        out.key[(size_t)e * out.m + s] = jc.key[(size_t)e * jc.m + c];
    }
}

// range[2e], range[2e+1] = min, max of key field e
__global__ void k_key_range_init(int32_t* __restrict__ range) {
    if (threadIdx.x < 14) {
        range[threadIdx.x] = (threadIdx.x & 1) ? INT_MIN : INT_MAX;
    }
}

__global__ __launch_bounds__(kBlock) void k_key_range(const int32_t* __restrict__ key, int m,
                                                      int32_t* __restrict__ range) {
    const int s = blockIdx.x * kBlock + threadIdx.x;
    for (int e = 0; e < 7; ++e) {
        int lo = INT_MAX, hi = INT_MIN;
        if (s < m) {
            lo = hi = key[(size_t)e * m + s];
        }
        for (int d = kWave / 2; d > 0; d >>= 1) {
            lo = min(lo, __shfl_xor(lo, d, kWave));
            hi = max(hi, __shfl_xor(hi, d, kWave));
        }
        if ((threadIdx.x & (kWave - 1)) == 0) {
            atomicMin(&range[2 * e], lo);
            atomicMax(&range[2 * e + 1], hi);
        }
    }
}

// gather one key field through a permutation, biased to be non-negative
__global__ __launch_bounds__(kBlock) void k_gather_field(const int32_t* __restrict__ field,
                                                         const int32_t* __restrict__ perm, int m, int32_t bias,
                                                         int32_t* __restrict__ out) {
    const int s = blockIdx.x * kBlock + threadIdx.x;
    if (s < m) {
        out[s] = field[perm[s]] + bias;
    }
}

// first[s] = 1 if sorted candidate s starts a new key (the kept one of its group)
__global__ __launch_bounds__(kBlock) void k_key_heads(const int32_t* __restrict__ key, const int32_t* __restrict__ perm,
                                                      int m, int32_t* __restrict__ head) {
    const int s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= m) {
        return;
    }
    bool h = s == 0;
    if (!h) {
        const int p = perm[s], q = perm[s - 1];
        for (int e = 0; e < 7 && !h; ++e) {
            h = key[(size_t)e * m + p] != key[(size_t)e * m + q];
        }
    }
    head[s] = h ? 1 : 0;
}

// The kept seeds in key order, and their cells in the drop hash
// (cell = 0.4 hmax, c3 = floor((x + v) / cell) clamped to [0, hd)).
struct KeptSeeds {
    float* x;        // [3 nk]
    int32_t* lab3;   // [3 nk]
    int32_t* bkey;   // [nk]
    int nk;
    float cell, ox, oy, oz;   // cell size, offsets (the voxel sizes)
    int hd[3];
};

__device__ __forceinline__ int drop_cell(float p, float o, float cell, int hd) {
    const int c = (int)floorf((p + o) / cell);
    return c < 0 ? 0 : (c > hd - 1 ? hd - 1 : c);
}

__global__ __launch_bounds__(kBlock) void k_keep_seeds(const JunctionCands jc, const int32_t* __restrict__ perm,
                                                       const int32_t* __restrict__ head_offs, KeptSeeds ks) {
    const int s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= jc.m || head_offs[s + 1] == head_offs[s]) {
        return;
    }
    const int c = perm[s], k = head_offs[s];
    const float x = jc.x[3 * (size_t)c], y = jc.x[3 * (size_t)c + 1], z = jc.x[3 * (size_t)c + 2];
    ks.x[3 * (size_t)k] = x;
    ks.x[3 * (size_t)k + 1] = y;
    ks.x[3 * (size_t)k + 2] = z;
    for (int e = 0; e < 3; ++e) {
        ks.lab3[3 * (size_t)k + e] = jc.lab3[3 * (size_t)c + e];
    }
    const int cx = drop_cell(x, ks.ox, ks.cell, ks.hd[0]), cy = drop_cell(y, ks.oy, ks.cell, ks.hd[1]),
              cz = drop_cell(z, ks.oz, ks.cell, ks.hd[2]);
    ks.bkey[k] = cx + ks.hd[0] * (cy + ks.hd[1] * cz);
}

// keep[i] = 0 for a lattice node closer than 0.4 h to a kept seed
// (src/v2m_particles.cpp:183-214); sorted: the seeds' bkeys ascending, with sidx
// the seed of each
__global__ __launch_bounds__(kBlock) void k_drop_near_seeds(Field f, int n, const float* __restrict__ P,
                                                            const uint8_t* __restrict__ typ, KeptSeeds ks,
                                                            const int32_t* __restrict__ sorted_bkey,
                                                            const int32_t* __restrict__ sidx,
                                                            int32_t* __restrict__ keep) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) {
        return;
    }
    keep[i] = 1;
    if (typ[i] == TN_CORNER) {
        return;
    }
    const float px = P[3 * (size_t)i], py = P[3 * (size_t)i + 1], pz = P[3 * (size_t)i + 2];
    const float h = size_at(f, px, py, pz);
    const int c0 = drop_cell(px, ks.ox, ks.cell, ks.hd[0]), c1 = drop_cell(py, ks.oy, ks.cell, ks.hd[1]),
              c2 = drop_cell(pz, ks.oz, ks.cell, ks.hd[2]);
    for (int dz = -1; dz <= 1; ++dz) {
        for (int dy = -1; dy <= 1; ++dy) {
            for (int dx = -1; dx <= 1; ++dx) {
                const int x = c0 + dx, y = c1 + dy, z = c2 + dz;
                if (x < 0 || y < 0 || z < 0 || x >= ks.hd[0] || y >= ks.hd[1] || z >= ks.hd[2]) {
                    continue;
                }
                const int kk = x + ks.hd[0] * (y + ks.hd[1] * z);
                int lo = 0, hi = ks.nk;   // lower bound of kk
                while (lo < hi) {
                    const int mid = (lo + hi) >> 1;
                    if (sorted_bkey[mid] < kk) {
                        lo = mid + 1;
                    } else {
                        hi = mid;
                    }
                }
                for (int t = lo; t < ks.nk && sorted_bkey[t] == kk; ++t) {
                    const int q = sidx[t];
                    const float ex = px - ks.x[3 * (size_t)q], ey = py - ks.x[3 * (size_t)q + 1],
                                ez = pz - ks.x[3 * (size_t)q + 2];
                    if (ex * ex + ey * ey + ez * ez < 0.16f * h * h) {
                        keep[i] = 0;
                        return;
                    }
                }
            }
        }
    }
}

// ---- the feature nodes of a shape input (add_feature_nodes, src/v2m_particles.cpp:
// 304-620). The host samples the curves and keeps the candidates; the device answers
// its label queries and drops the seeds near the kept ones. ----------------------------

// the labels round each point, as bits: the point and 6 samples at 0.2 h
// (add_feature_nodes' mask_at, src/v2m_particles.cpp:312-326)
__global__ __launch_bounds__(kBlock) void k_feat_mask(Field f, const float* __restrict__ X, int m,
                                                      unsigned long long* __restrict__ mask) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= m) {
        return;
    }
    const float x = X[3 * (size_t)i], y = X[3 * (size_t)i + 1], z = X[3 * (size_t)i + 2];
    const float dir[7][3] = { { 0, 0, 0 }, { 1, 0, 0 }, { -1, 0, 0 }, { 0, 1, 0 }, { 0, -1, 0 }, { 0, 0, 1 }, { 0, 0, -1 } };
    const float e = 0.2f * size_at(f, x, y, z);
    unsigned long long bits = 0;
    for (int s = 0; s < 7; ++s) {
        const int l = owner_at(f, p3(x + e * dir[s][0], y + e * dir[s][1], z + e * dir[s][2])).x;
        bits |= l < 64 ? 1ull << l : 0ull;
    }
    mask[i] = bits;
}

// the label at each of nd samples at 0.25 h round each candidate
// (src/v2m_particles.cpp:482-494); one thread per (candidate, direction)
__global__ __launch_bounds__(kBlock) void k_feat_labels(Field f, const float* __restrict__ X, int nc,
                                                        const float* __restrict__ dirs, int nd,
                                                        int32_t* __restrict__ lab) {
    const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (t >= (int64_t)nc * nd) {
        return;
    }
    const int c = (int)(t / nd), d = (int)(t % nd);
    const float x = X[3 * (size_t)c], y = X[3 * (size_t)c + 1], z = X[3 * (size_t)c + 2];
    const float e = 0.25f * size_at(f, x, y, z);
    lab[t] = owner_at(f, p3(x + e * dirs[3 * d], y + e * dirs[3 * d + 1], z + e * dirs[3 * d + 2])).x;
}

// keep[i] = 0 for a seed closer than 0.4 h to a feature node (src/v2m_particles.cpp:
// 563-589; the reference's hash only narrows the search to the cells within reach)
__global__ __launch_bounds__(kBlock) void k_feat_drop(Field f, int n, const float* __restrict__ P,
                                                      const float* __restrict__ K, int nk,
                                                      int32_t* __restrict__ keep) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) {
        return;
    }
    const float x = P[3 * (size_t)i], y = P[3 * (size_t)i + 1], z = P[3 * (size_t)i + 2];
    const float r = 0.4f * size_at(f, x, y, z);
    int k = 1;
    for (int j = 0; j < nk && k; ++j) {
        const float dx = K[3 * (size_t)j] - x, dy = K[3 * (size_t)j + 1] - y, dz = K[3 * (size_t)j + 2] - z;
        k = dx * dx + dy * dy + dz * dz < r * r ? 0 : 1;
    }
    keep[i] = k;
}

// The node arrays of a seeding (part3: the third partner of a corner node).
struct SeedNodes {
    float* P;
    uint16_t* lab;
    uint8_t* typ;
    uint16_t* part;
    uint16_t* part3;
};

// the surviving lattice nodes at their scanned places, then the kept seeds as
// JUNCTION nodes (own, partners o0 o1) after them
__global__ __launch_bounds__(kBlock) void k_merge_junctions(SeedNodes in, int n, const int32_t* __restrict__ keep_offs,
                                                            KeptSeeds ks, SeedNodes out) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    const int n1 = keep_offs[n];
    if (i < n) {
        if (keep_offs[i + 1] != keep_offs[i]) {
            const int t = keep_offs[i];
            out.P[3 * (size_t)t] = in.P[3 * (size_t)i];
            out.P[3 * (size_t)t + 1] = in.P[3 * (size_t)i + 1];
            out.P[3 * (size_t)t + 2] = in.P[3 * (size_t)i + 2];
            out.lab[t] = in.lab[i];
            out.typ[t] = in.typ[i];
            out.part[2 * (size_t)t] = in.part[2 * (size_t)i];
            out.part[2 * (size_t)t + 1] = in.part[2 * (size_t)i + 1];
            out.part3[t] = in.part3[i];
        }
    } else if (i < n + ks.nk) {
        const int k = i - n, t = n1 + k;
        out.P[3 * (size_t)t] = ks.x[3 * (size_t)k];
        out.P[3 * (size_t)t + 1] = ks.x[3 * (size_t)k + 1];
        out.P[3 * (size_t)t + 2] = ks.x[3 * (size_t)k + 2];
        out.lab[t] = (uint16_t)ks.lab3[3 * (size_t)k];
        out.typ[t] = TN_JUNCTION;
        out.part[2 * (size_t)t] = (uint16_t)ks.lab3[3 * (size_t)k + 1];
        out.part[2 * (size_t)t + 1] = (uint16_t)ks.lab3[3 * (size_t)k + 2];
        out.part3[t] = kNoLab;
    }
}

__device__ __forceinline__ float det3(const float M[3][3]) {
    return M[0][0] * (M[1][1] * M[2][2] - M[1][2] * M[2][1]) - M[0][1] * (M[1][0] * M[2][2] - M[1][2] * M[2][0]) +
           M[0][2] * (M[1][0] * M[2][1] - M[1][1] * M[2][0]);
}

// v2m_corner_vertex's placement, src/opencl/v2m_seed_body.cl:656-729, for corner
// vertex s of the list, written as node out0 + s
__global__ __launch_bounds__(kBlock) void k_corner_fill(Field f, VertexGrid vg, const int32_t* __restrict__ list, int m,
                                                        SeedNodes out, int out0) {
    const int s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= m) {
        return;
    }
    int i, j, k;
    vertex_ijk(vg, list[s], i, j, k);
    int labs[8];
    const int n = vertex_labels(f, i, j, k, labs);
    P3 x = p3((i + 0.5f) * f.vx, (j + 0.5f) * f.vy, (k + 0.5f) * f.vz);
    float ph[8];
    for (int t = 0; t < n; ++t) {
        ph[t] = phi_point(f, labs[t], x);
    }
    for (int t = 0; t < n; ++t) {   // by phi, descending (exchange sort)
        for (int u = t + 1; u < n; ++u) {
            if (ph[u] > ph[t]) {
                const float w = ph[t];
                ph[t] = ph[u];
                ph[u] = w;
                const int l = labs[t];
                labs[t] = labs[u];
                labs[u] = l;
            }
        }
    }
    int ia = 0;
    while (ia < 4 && labs[ia] == 0) {
        ++ia;
    }
    const int a = labs[ia];
    int o[3], no = 0;
    for (int t = 0; t < 4; ++t) {
        if (t != ia) {
            o[no++] = labs[t];
        }
    }
    // Newton on (psi_a o0, psi_a o1, psi_a o2) = 0, the steps by Cramer's rule
    float q[3] = { x.x, x.y, x.z };
    bool ok = true;
    for (int it = 0; it < 5 && ok; ++it) {
        float J[3][3], v[3];
        for (int r = 0; r < 3; ++r) {
            const float4 g = psi_grad_point(f.self, a, o[r], p3(q[0], q[1], q[2]));
            J[r][0] = g.x;
            J[r][1] = g.y;
            J[r][2] = g.z;
            v[r] = g.w;
        }
        const float det = det3(J);
        if (fabsf(det) < 1e-12f) {
            ok = false;
            break;
        }
        float dlt[3];
        for (int c = 0; c < 3; ++c) {
            float M[3][3];
            for (int r = 0; r < 3; ++r) {
                for (int cc = 0; cc < 3; ++cc) {
                    M[r][cc] = cc == c ? -v[r] : J[r][cc];
                }
            }
            dlt[c] = det3(M) / det;
        }
        for (int c = 0; c < 3; ++c) {
            q[c] += dlt[c];
        }
    }
    const float ex = over_vx(f, q[0] - x.x), ey = over_vy(f, q[1] - x.y), ez = over_vz(f, q[2] - x.z);
    if (ok && ex * ex + ey * ey + ez * ez <= 0.25f) {
        x = p3(q[0], q[1], q[2]);
    }
    const size_t t = (size_t)out0 + s;
    out.P[3 * t] = x.x;
    out.P[3 * t + 1] = x.y;
    out.P[3 * t + 2] = x.z;
    out.lab[t] = (uint16_t)a;
    out.typ[t] = TN_CORNER;
    out.part[2 * t] = (uint16_t)o[0];
    out.part[2 * t + 1] = (uint16_t)o[1];
    out.part3[t] = (uint16_t)o[2];
}

}  // namespace tnk
