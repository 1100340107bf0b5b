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
};

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
                const int l = owner_at(f, p).x;
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
