// tn_device.hpp -- per-node device bodies of the relaxation for gfx950.
//
// Field queries on the brick-sparse interface fields, the multi-level hash,
// the K-nearest truss list, the bar forces and the trapped move. Each function
// names the reference lines whose arithmetic it reproduces: the results are
// compared bit for bit with the CPU path, so every expression keeps the
// reference's association and operand order, and the library is compiled with
// -ffp-contract=off, correctly rounded fp32 division and square root, and fp32
// denormals kept (the CPU path's IEEE single precision).
//
// The interface queries of the move (psi, projections, label tests) are real
// calls (__noinline__) only on their rare paths (junction projections, third
// labels, cells across bricks), taking and returning points by value; the hot
// paths are inline and share trilinear-cell caches (PsiCell).
//
// Layout in HBM (device buffers of a plan), by node slot t (nodes renumbered in cell
// order at every rebuild; id[t] is the node's index in the caller's arrays):
//   X[t]    float4 (x, y, z, h at the node)  -- positions and sizes, one load per gather
//   typ[t]  u8, lab[t] u16, part[t] u32 (partner b | partner c << 16)
//   nbrT[k*n + t]  int32 slots, list-position-major K lists (coalesced), nnb[t] int32
//   Ps[t]   float4 positions and sizes as of the last rebuild (the search's candidates)
#pragma once

#include <hip/hip_runtime.h>
#include <limits.h>
#include <stdint.h>

#include "../../include/trussnet_amd.h"

#pragma clang fp contract(off)

namespace tnd {

constexpr int kK = TN_K;
constexpr int kBL = TN_BL;
constexpr int kNoLab = TN_NOLAB;

// The grid on the device (a tn_grid_view with device pointers).
struct Field {
    int nx, ny, nz, nbx, nby, nbz;
    float vx, vy, vz;
    float ivx, ivy, ivz;   // 1 / v, used only when every voxel size is a power of two
    int pow2;
    const uint16_t* __restrict__ L;
    const int32_t* __restrict__ bl_cnt;
    const uint16_t* __restrict__ bl_lab;
    const int32_t* __restrict__ bl_slot;
    const float* __restrict__ phi;
    const float* __restrict__ h;
    const float* __restrict__ gI;    // gray-scale mode: the intensity per voxel
    const float* __restrict__ gTW;   // gray-scale mode: [t_0..t_{m-1}, W_0..W_{m-1}]
    int gm;                          // gray-scale mode: m > 0 thresholds (0: a label volume)
    // this struct's copy in device memory. Out-of-line functions take it instead of a
    // reference: a reference to the kernel argument would move the whole struct into
    // scratch memory and turn the kernel's loads through its pointers into flat loads.
    const struct Field* self;
};

// The multi-level hash geometry (V2mHash, src/opencl/v2m_particle_body.cl:21-27).
struct Hash {
    float b0;
    int nlev;
    float ox, oy, oz;
    int dim[TN_MAXLEV][3];
    int off[TN_MAXLEV];
    int nkeys;
    int prune;   // every node's reach fits its level's cell: the K-nearest search may skip far cells
};

// x / v for the voxel sizes: when they are powers of two, x * (1 / v) is the same
// correctly rounded value (both round the exact x 2^-k), and a multiplication is a
// fraction of an IEEE division's cost
__device__ __forceinline__ float over_vx(const Field& f, float x) {
    return f.pow2 ? x * f.ivx : x / f.vx;
}
__device__ __forceinline__ float over_vy(const Field& f, float x) {
    return f.pow2 ? x * f.ivy : x / f.vy;
}
__device__ __forceinline__ float over_vz(const Field& f, float x) {
    return f.pow2 ? x * f.ivz : x / f.vz;
}

struct P3 {
    float x, y, z;
};

// a projected point and whether the projection succeeded
struct P3ok {
    P3 p;
    int ok;
};

__device__ __forceinline__ P3 p3(float x, float y, float z) {
    P3 r;
    r.x = x;
    r.y = y;
    r.z = z;
    return r;
}

__device__ __forceinline__ size_t vox(const Field& f, int i, int j, int k) {
    return (size_t)i + (size_t)f.nx * ((size_t)j + (size_t)f.ny * (size_t)k);
}

// v2m_brick_of, src/opencl/v2m_grid_body.cl:116-118
__device__ __forceinline__ int brick_at(const Field& f, int i, int j, int k) {
    return (i >> 3) + f.nbx * ((j >> 3) + f.nby * (k >> 3));
}

// A brick's label list (bl_lab[16 b .. 16 b + 16): ascending, padded), its count
// and first slot, read with two 16-byte loads and two scalar ones at once rather
// than label by label.
struct BrickList {
    uint4 lo, hi;
    int cnt, s0;
};

__device__ __forceinline__ BrickList brick_list(const Field& f, int b) {
    const uint4* w = reinterpret_cast<const uint4*>(f.bl_lab + (size_t)b * kBL);
    BrickList r;
    r.lo = w[0];
    r.hi = w[1];
    r.cnt = f.bl_cnt[b];
    r.s0 = f.bl_slot[b];
    return r;
}

// the first position s < min(cnt, 16) of the list holding label l, or -1
__device__ __forceinline__ int list_find(const BrickList& bl, int l) {
    const uint32_t w[8] = { bl.lo.x, bl.lo.y, bl.lo.z, bl.lo.w, bl.hi.x, bl.hi.y, bl.hi.z, bl.hi.w };
    const int n = bl.cnt < kBL ? bl.cnt : kBL;
    int pos = -1;
#pragma unroll
    for (int s = kBL - 1; s >= 0; --s) {
        const int v = (int)((w[s >> 1] >> ((s & 1) * 16)) & 0xFFFFu);
        pos = (s < n && v == l) ? s : pos;
    }
    return pos;
}

// v2m_slot_of, src/opencl/v2m_grid_body.cl:122-134, on a loaded list
__device__ __forceinline__ int list_slot(const BrickList& bl, int l, int* has) {
    const int pos = list_find(bl, l);
    *has = pos >= 0;
    return (pos < 0 || bl.s0 < 0) ? -1 : bl.s0 + pos;
}

__device__ __forceinline__ int slot_in(const Field& f, int b, int l, int* has) {
    return list_slot(brick_list(f, b), l, has);
}

// v2m_phi_vox, src/opencl/v2m_grid_body.cl:137-153
__device__ __forceinline__ float phi_voxel(const Field& f, int l, int i, int j, int k) {
    if (i < 0 || j < 0 || k < 0 || i >= f.nx || j >= f.ny || k >= f.nz) {
        return l == 0 ? 1.0f : 0.0f;
    }
    int has;
    const int s = slot_in(f, brick_at(f, i, j, k), l, &has);
    if (s < 0) {
        return (has && f.L[vox(f, i, j, k)] == l) ? 1.0f : 0.0f;
    }
    return f.phi[(size_t)s * TN_SLOT + (size_t)((i & 7) + TN_BS * ((j & 7) + TN_BS * (k & 7)))];
}

// v2m_phi_cell's single-brick case (src/opencl/v2m_grid_body.cl:161-180): the cell
// (i0,j0,k0) lies in one brick inside the grid
__device__ __forceinline__ bool cell_in_one_brick(const Field& f, int i0, int j0, int k0) {
    return i0 >= 0 && j0 >= 0 && k0 >= 0 && i0 + 1 < f.nx && j0 + 1 < f.ny && k0 + 1 < f.nz &&
           (i0 >> 3) == ((i0 + 1) >> 3) && (j0 >> 3) == ((j0 + 1) >> 3) && (k0 >> 3) == ((k0 + 1) >> 3);
}

// the 8 corners of phi_l in a single-brick cell, given l's slot s in that brick
// (s < 0: uniform or absent, has = l among the brick's labels)
__device__ __forceinline__ void corners_in_brick(const Field& f, int l, int s, int has, int i0, int j0, int k0,
                                                 float* c) {
    if (s >= 0) {
        const float* base = f.phi + (size_t)s * TN_SLOT + (size_t)((i0 & 7) + TN_BS * ((j0 & 7) + TN_BS * (k0 & 7)));
#pragma unroll
        for (int m = 0; m < 8; ++m) {
            c[m] = base[(m & 1) + TN_BS * ((m >> 1) & 1) + TN_BS * TN_BS * (m >> 2)];
        }
    } else {
#pragma unroll
        for (int m = 0; m < 8; ++m) {
            c[m] = (has && f.L[vox(f, i0 + (m & 1), j0 + ((m >> 1) & 1), k0 + (m >> 2))] == l) ? 1.0f : 0.0f;
        }
    }
}

// v2m_phi_cell's general case (src/opencl/v2m_grid_body.cl:182-185): a cell across
// bricks or at the grid's edge, corner by corner as v2m_phi_vox reads each. The
// corners lie in at most eight bricks, mostly two or four; each distinct brick's
// label list is read once (for both labels of a pair).
struct Corners {
    float v[8];
};

struct PairCorners {
    float a[8], b[8];
};

template <bool kPair>
__device__ __forceinline__ void corners_split(const Field& f, int la, int lb, int i0, int j0, int k0, float* ca,
                                              float* cb) {
    // the slots of the last two distinct bricks met (a cell across one brick face
    // alternates between two)
    int b0 = -1, b1 = -1, sa0 = -1, sb0 = -1, ha0 = 0, hb0 = 0, sa1 = -1, sb1 = -1, ha1 = 0, hb1 = 0;
#pragma unroll
    for (int m = 0; m < 8; ++m) {
        const int i = i0 + (m & 1), j = j0 + ((m >> 1) & 1), k = k0 + (m >> 2);
        if (i < 0 || j < 0 || k < 0 || i >= f.nx || j >= f.ny || k >= f.nz) {   // outside the grid: label 0
            ca[m] = la == 0 ? 1.0f : 0.0f;
            if (kPair) {
                cb[m] = lb == 0 ? 1.0f : 0.0f;
            }
            continue;
        }
        const int bk = brick_at(f, i, j, k);
        if (bk != b0) {
            if (bk == b1) {   // swap the two entries
                int t = b0; b0 = b1; b1 = t;
                t = sa0; sa0 = sa1; sa1 = t;
                t = sb0; sb0 = sb1; sb1 = t;
                t = ha0; ha0 = ha1; ha1 = t;
                t = hb0; hb0 = hb1; hb1 = t;
            } else {
                b1 = b0;
                sa1 = sa0;
                sb1 = sb0;
                ha1 = ha0;
                hb1 = hb0;
                const BrickList bl = brick_list(f, bk);
                b0 = bk;
                sa0 = list_slot(bl, la, &ha0);
                if (kPair) {
                    sb0 = list_slot(bl, lb, &hb0);
                }
            }
        }
        const size_t off = (size_t)((i & 7) + TN_BS * ((j & 7) + TN_BS * (k & 7)));
        const int lv = (sa0 < 0 || (kPair && sb0 < 0)) ? (int)f.L[vox(f, i, j, k)] : -1;
        ca[m] = sa0 >= 0 ? f.phi[(size_t)sa0 * TN_SLOT + off] : ((ha0 && lv == la) ? 1.0f : 0.0f);
        if (kPair) {
            cb[m] = sb0 >= 0 ? f.phi[(size_t)sb0 * TN_SLOT + off] : ((hb0 && lv == lb) ? 1.0f : 0.0f);
        }
    }
}

__device__ __noinline__ Corners phi_corners_split(const Field* fp, int l, int i0, int j0, int k0) {
    const Field f = *fp;
    Corners c;
    corners_split<false>(f, l, -1, i0, j0, k0, c.v, nullptr);
    return c;
}

__device__ __noinline__ PairCorners pair_corners_split(const Field* fp, int a, int b, int i0, int j0, int k0) {
    const Field f = *fp;
    PairCorners c;
    corners_split<true>(f, a, b, i0, j0, k0, c.a, c.b);
    return c;
}

// v2m_phi_cell, src/opencl/v2m_grid_body.cl:159-186 (corner bit 0 -> +i, 1 -> +j, 2 -> +k)
__device__ __forceinline__ void phi_corners(const Field& f, int l, int i0, int j0, int k0, float* c) {
    if (cell_in_one_brick(f, i0, j0, k0)) {
        int has;
        const int s = slot_in(f, brick_at(f, i0, j0, k0), l, &has);
        corners_in_brick(f, l, s, has, i0, j0, k0, c);
        return;
    }
    const Corners cc = phi_corners_split(f.self, l, i0, j0, k0);
#pragma unroll
    for (int m = 0; m < 8; ++m) {
        c[m] = cc.v[m];
    }
}

// A trilinear cell of psi_ab = phi_a - phi_b kept across evaluations of one label
// pair: a bracketing or a crossing search probes a short segment, mostly within
// one cell and one brick, so the corner differences (and the pair's slots in the
// brick) are loaded once. The values are the ones v2m_psi computes, so every
// result is unchanged.
struct PsiCell {
    int i0, j0, k0;           // the cell held in d (i0 = INT_MIN: none)
    int bk, sa, sb, ha, hb;   // the brick whose slots for a and b are known (bk = -1: none)
    float d[8];               // phi_a - phi_b at the 8 corners
};

__device__ __forceinline__ void psi_cell_init(PsiCell& pc) {
    pc.i0 = INT_MIN;
    pc.j0 = pc.k0 = 0;
    pc.bk = -1;
    pc.sa = pc.sb = -1;
    pc.ha = pc.hb = 0;
}

__device__ __forceinline__ void psi_cell_load(const Field& f, int a, int b, int i0, int j0, int k0, PsiCell& pc) {
    if (i0 == pc.i0 && j0 == pc.j0 && k0 == pc.k0) {
        return;
    }
    pc.i0 = i0;
    pc.j0 = j0;
    pc.k0 = k0;
    float c[8], e[8];
    if (cell_in_one_brick(f, i0, j0, k0)) {
        const int bk = brick_at(f, i0, j0, k0);
        if (bk != pc.bk) {
            pc.bk = bk;
            const BrickList bl = brick_list(f, bk);
            pc.sa = list_slot(bl, a, &pc.ha);
            pc.sb = list_slot(bl, b, &pc.hb);
        }
        corners_in_brick(f, a, pc.sa, pc.ha, i0, j0, k0, c);
        corners_in_brick(f, b, pc.sb, pc.hb, i0, j0, k0, e);
    } else {
        const PairCorners pc2 = pair_corners_split(f.self, a, b, i0, j0, k0);
#pragma unroll
        for (int m = 0; m < 8; ++m) {
            c[m] = pc2.a[m];
            e[m] = pc2.b[m];
        }
    }
#pragma unroll
    for (int m = 0; m < 8; ++m) {
        pc.d[m] = c[m] - e[m];
    }
}

// the trilinear tail of v2m_phi_at / v2m_psi (src/opencl/v2m_grid_body.cl:196-199)
__device__ __forceinline__ float trilerp(const float* c, float fx, float fy, float fz) {
    const float x00 = c[0] + fx * (c[1] - c[0]), x10 = c[2] + fx * (c[3] - c[2]);
    const float x01 = c[4] + fx * (c[5] - c[4]), x11 = c[6] + fx * (c[7] - c[6]);
    const float y0 = x00 + fy * (x10 - x00), y1 = x01 + fy * (x11 - x01);
    return y0 + fz * (y1 - y0);
}

// ---- shape fields (gm < 0): the labels' analytic-field program ----------------------
// The program layout and op codes are v2mesh's (src/opencl/v2m_sdf_body.cl:1-45):
// [0] labels N, [1] field width w, [2..4] origin, [5..5+N-1] code offsets, [5+N] crease
// blend radius, [6+N] gap-closing radius, [7+N] cull margin, [8+N] brick table
// offset (0: none); code: an op, its operands; op 0 ends a label's code.
namespace sdf {
constexpr int kEnd = 0, kPrim = 1, kMax = 2, kMin = 3, kNeg = 4, kConst = 5, kBBox = 6, kMConst = 7, kAdd = 8,
              kScale = 9, kCMax = 10, kCMin = 11;
constexpr int kSphere = 1, kBox = 2, kCyl = 3, kPlane = 4, kSlab = 5, kTorus = 6, kEllipsoid = 7, kCone = 8,
              kInfCyl = 9;
constexpr int kStack = 12;

__host__ __device__ __forceinline__ int nparam(int type) {
    return type == kSphere ? 4
           : (type == kBox || type == kPlane) ? 6
           : (type == kCyl || type == kInfCyl) ? 7
           : (type == kSlab || type == kTorus || type == kCone) ? 8
           : type == kEllipsoid ? 15 : 0;
}

// one primitive's signed distance, > 0 inside (v2m_sdf_prim, :76-161)
__device__ float prim(int type, const float* __restrict__ q, float x, float y, float z) {
    if (type == kSphere) {
        const float dx = x - q[0], dy = y - q[1], dz = z - q[2];
        return q[3] - sqrtf(dx * dx + dy * dy + dz * dz);
    }
    if (type == kBox) {
        const float cx = 0.5f * (q[0] + q[3]), cy = 0.5f * (q[1] + q[4]), cz = 0.5f * (q[2] + q[5]);
        const float ux = fabsf(x - cx) - 0.5f * fabsf(q[3] - q[0]), uy = fabsf(y - cy) - 0.5f * fabsf(q[4] - q[1]),
                    uz = fabsf(z - cz) - 0.5f * fabsf(q[5] - q[2]);
        const float ox = fmaxf(ux, 0.0f), oy = fmaxf(uy, 0.0f), oz = fmaxf(uz, 0.0f);
        const float out = sqrtf(ox * ox + oy * oy + oz * oz), in = fminf(fmaxf(ux, fmaxf(uy, uz)), 0.0f);
        return -(out + in);
    }
    if (type == kPlane) {
        return -((x - q[0]) * q[3] + (y - q[1]) * q[4] + (z - q[2]) * q[5]);
    }
    if (type == kSlab) {
        const float t = (x - q[0]) * q[3] + (y - q[1]) * q[4] + (z - q[2]) * q[5];
        return fminf(t - q[6], q[7] - t);
    }
    if (type == kInfCyl) {
        const float dx = x - q[0], dy = y - q[1], dz = z - q[2];
        const float t = dx * q[3] + dy * q[4] + dz * q[5];
        const float rx = dx - t * q[3], ry = dy - t * q[4], rz = dz - t * q[5];
        return q[6] - sqrtf(rx * rx + ry * ry + rz * rz);
    }
    if (type == kTorus) {
        const float dx = x - q[0], dy = y - q[1], dz = z - q[2];
        const float hh = dx * q[3] + dy * q[4] + dz * q[5];
        const float rx = dx - hh * q[3], ry = dy - hh * q[4], rz = dz - hh * q[5];
        const float a = sqrtf(rx * rx + ry * ry + rz * rz) - q[6];
        return q[7] - sqrtf(a * a + hh * hh);
    }
    if (type == kCyl) {   // capped cylinder, I. Quilez's exact form
        const float bx = q[3] - q[0], by = q[4] - q[1], bz = q[5] - q[2];
        const float px = x - q[0], py = y - q[1], pz = z - q[2];
        const float baba = bx * bx + by * by + bz * bz, paba = px * bx + py * by + pz * bz;
        const float ex = px * baba - bx * paba, ey = py * baba - by * paba, ez = pz * baba - bz * paba;
        const float u = sqrtf(ex * ex + ey * ey + ez * ez) - q[6] * baba;
        const float v = fabsf(paba - baba * 0.5f) - baba * 0.5f;
        const float u2 = u * u, v2 = v * v * baba;
        const float dd = fmaxf(u, v) < 0.0f ? -fminf(u2, v2) : ((u > 0.0f ? u2 : 0.0f) + (v > 0.0f ? v2 : 0.0f));
        return -(dd < 0.0f ? -1.0f : 1.0f) * sqrtf(fabsf(dd)) / baba;
    }
    if (type == kCone) {   // capped frustum, I. Quilez's exact form
        const float bx = q[3] - q[0], by = q[4] - q[1], bz = q[5] - q[2];
        const float px = x - q[0], py = y - q[1], pz = z - q[2];
        const float ra = q[6], rb = q[7];
        const float rba = rb - ra, baba = bx * bx + by * by + bz * bz;
        const float papa = px * px + py * py + pz * pz, paba = (px * bx + py * by + pz * bz) / baba;
        const float xr = sqrtf(fmaxf(papa - paba * paba * baba, 0.0f));
        const float cax = fmaxf(0.0f, xr - ((paba < 0.5f) ? ra : rb)), cay = fabsf(paba - 0.5f) - 0.5f;
        const float k = rba * rba + baba;
        const float fr = fminf(fmaxf((rba * (xr - ra) + paba * baba) / k, 0.0f), 1.0f);
        const float cbx = xr - ra - fr * rba, cby = paba - fr;
        const float sg = (cbx < 0.0f && cay < 0.0f) ? -1.0f : 1.0f;
        const float d2 = fminf(cax * cax + cay * cay * baba, cbx * cbx + cby * cby * baba);
        return -sg * sqrtf(d2);
    }
    if (type == kEllipsoid) {   // the bound k0 (k0 - 1) / k1, exact on the surface
        const float dx = x - q[0], dy = y - q[1], dz = z - q[2];
        const float lx = q[6] * dx + q[7] * dy + q[8] * dz, ly = q[9] * dx + q[10] * dy + q[11] * dz,
                    lz = q[12] * dx + q[13] * dy + q[14] * dz;
        const float ax = lx / q[3], ay = ly / q[4], az = lz / q[5];
        const float bx = ax / q[3], by = ay / q[4], bz = az / q[5];
        const float k0 = sqrtf(ax * ax + ay * ay + az * az), k1 = sqrtf(bx * bx + by * by + bz * bz);
        if (k1 < 1e-20f) {
            return fminf(q[3], fminf(q[4], q[5]));
        }
        return -k0 * (k0 - 1.0f) / k1;
    }
    return -1e30f;
}

// (value, gradient): .w the value
__device__ __forceinline__ float4 sv(float v) {
    return make_float4(0.0f, 0.0f, 0.0f, v);
}

// min (sgn 1) or max (sgn -1) of a and b with the crease blend of radius k and the
// gap closing of radius kgap (v2m_sdf_minmax, :191-273)
__device__ float4 minmax(float4 a, float4 b, float sgn, float k, float kgap, int grad, int* near) {
    const float va = sgn * a.w, vb = sgn * b.w;
    const bool sliver = kgap > 0.0f && va > -kgap && vb > -kgap && va + vb < kgap;
    if (sliver && !grad) {
        *near = 1;
    }
    float pen = 0.0f, dpen = 0.0f;
    if (sliver && grad) {
        const float ga = sqrtf(a.x * a.x + a.y * a.y + a.z * a.z), gb = sqrtf(b.x * b.x + b.y * b.y + b.z * b.z);
        if (ga > 0.0f && gb > 0.0f) {
            const float c = (a.x * b.x + a.y * b.y + a.z * b.z) / (ga * gb);
            const float wg = fminf(1.0f, fmaxf(0.0f, (-c - 0.7f) / 0.25f));
            const float u = fmaxf(0.0f, 1.0f - (va + vb) / kgap);
            pen = kgap * wg * u * u;
            dpen = 2.0f * wg * u;
        }
    }
    float4 r = a;
    if (k > 0.0f && fabsf(va - vb) < k) {
        if (!grad) {
            *near = 1;
        } else {
            const float ga = sqrtf(a.x * a.x + a.y * a.y + a.z * a.z), gb = sqrtf(b.x * b.x + b.y * b.y + b.z * b.z);
            const float c = ga > 0.0f && gb > 0.0f ? (a.x * b.x + a.y * b.y + a.z * b.z) / (ga * gb) : 0.0f;
            k *= fminf(1.0f, fmaxf(0.0f, (0.995f - c) / 0.095f));
        }
    }
    if (k > 0.0f && fabsf(va - vb) < k) {
        const float h = 0.5f + 0.5f * (vb - va) / k;
        r.w = sgn * (vb * (1.0f - h) + va * h - k * h * (1.0f - h));
        r.x = h * a.x + (1.0f - h) * b.x;
        r.y = h * a.y + (1.0f - h) * b.y;
        r.z = h * a.z + (1.0f - h) * b.z;
    } else if (vb < va) {
        r = b;
    }
    r.w -= sgn * pen;
    if (dpen > 0.0f) {
        r.x += dpen * (a.x + b.x);
        r.y += dpen * (a.y + b.y);
        r.z += dpen * (a.z + b.z);
    }
    return r;
}

__host__ __device__ __forceinline__ bool is_minmax(int op) {
    return op == kMax || op == kMin || op == kCMax || op == kCMin;
}

// Label l's field at grid-mm point p and (wantg) its gradient, (gx, gy, gz, s)
// (v2m_sdf_evalv, :275-478): the postfix code of the point's brick on a stack whose
// top four entries live in registers (t0 the top) and the rest in a small spill
// array, values first and the gradients only where a gap or crease needs them.
__device__ __noinline__ float4 evalv(const float* __restrict__ prog, int l, float px, float py, float pz, int wantg) {
    const int N = (int)prog[0];
    if (l < 0 || l >= N) {
        return sv(-1e30f);
    }
    const float x = px + prog[2], y = py + prog[3], z = pz + prog[4];
    int start = (int)prog[5 + l];
    const int T = (int)prog[8 + N];
    if (T > 0) {
        const int nbx = (int)prog[T], nby = (int)prog[T + 1], nbz = (int)prog[T + 2];
        const float vx = prog[T + 3];
        int bi = (int)floorf(px / vx + 0.5f) >> 3, bj = (int)floorf(py / vx + 0.5f) >> 3,
            bk = (int)floorf(pz / vx + 0.5f) >> 3;
        bi = bi < 0 ? 0 : (bi >= nbx ? nbx - 1 : bi);
        bj = bj < 0 ? 0 : (bj >= nby ? nby - 1 : bj);
        bk = bk < 0 ? 0 : (bk >= nbz ? nbz - 1 : bk);
        start = (int)prog[(int)prog[T + 4 + bi + nbx * (bj + nby * bk)] + l];
    }
    const float kgap = prog[6 + N], kblend = prog[5 + N], margin = prog[7 + N];
    float4 t0 = sv(0.0f), t1 = t0, t2 = t0, t3 = t0, spill[kStack - 4];
    int sp = 0;
    int grad = wantg != 0, near = 0;
    for (int pass = 0; pass < 2; ++pass) {
        sp = 0;
        int pc = start;
        near = 0;
        for (int guard = 0; guard < 100000; ++guard) {
            const int op = (int)prog[pc];
            if (op == kEnd) {
                break;
            }
            if (is_minmax(op)) {   // t1 op t0 -> t0, pop
                if (sp > 1) {
                    t0 = minmax(t1, t0, (op == kMax || op == kCMax) ? -1.0f : 1.0f, kblend, op >= kCMax ? 0.0f : kgap, grad,
                                &near);
                    t1 = t2;
                    t2 = t3;
                    if (sp > 4) {
                        t3 = spill[sp - 5];
                    }
                    --sp;
                }
                pc += 1;
                continue;
            }
            if (op == kNeg) {
                t0 = make_float4(-t0.x, -t0.y, -t0.z, -t0.w);
                pc += 1;
                continue;
            }
            if (op == kAdd) {   // t1 + t0 -> t0, pop
                if (sp > 1) {
                    t0.w += t1.w;
                    t0.x += t1.x;
                    t0.y += t1.y;
                    t0.z += t1.z;
                    t1 = t2;
                    t2 = t3;
                    if (sp > 4) {
                        t3 = spill[sp - 5];
                    }
                    --sp;
                }
                pc += 1;
                continue;
            }
            if (op == kScale) {
                const float fs = prog[pc + 1];
                t0 = make_float4(t0.x * fs, t0.y * fs, t0.z * fs, t0.w * fs);
                pc += 2;
                continue;
            }
            float4 nv = sv(0.0f);
            if (op == kPrim) {   // the value and (grad) this primitive's central differences
                const int type = (int)prog[pc + 1];
                const float* q = prog + pc + 2;
                const float e = fmaxf(1e-4f * (fabsf(q[0]) + fabsf(q[1]) + fabsf(q[2]) + 1.0f), 1e-5f);
                const int ne = grad ? 7 : 1;
                float fv[7];
                for (int k = 0; k < ne; ++k) {
                    const float dd = k == 0 ? 0.0f : ((k & 1) ? e : -e);
                    const int ax = (k - 1) >> 1;
                    fv[k] = prim(type, q, x + (ax == 0 ? dd : 0.0f), y + (ax == 1 ? dd : 0.0f), z + (ax == 2 ? dd : 0.0f));
                }
                nv.w = fv[0];
                if (grad) {
                    nv.x = (fv[1] - fv[2]) / (2.0f * e);
                    nv.y = (fv[3] - fv[4]) / (2.0f * e);
                    nv.z = (fv[5] - fv[6]) / (2.0f * e);
                }
                pc += 2 + nparam(type);
            } else if (op == kBBox) {   // outside the box (+ margin): the bound, its code skipped
                const float* q = prog + pc + 1;
                if (!(x < q[0] - margin || y < q[1] - margin || z < q[2] - margin || x > q[3] + margin ||
                      y > q[4] + margin || z > q[5] + margin)) {
                    pc += 9;
                    continue;
                }
                nv.w = q[7] * margin;
                pc += 9 + (int)q[6];
            } else if (op == kMConst) {
                nv.w = prog[pc + 1] * margin;
                pc += 2;
            } else if (op == kConst) {
                nv.w = prog[pc + 1];
                pc += 2;
            } else {
                break;   // a malformed program
            }
            const int op1 = (int)prog[pc], neg = op1 == kNeg;
            const int op2 = neg ? (int)prog[pc + 1] : op1;
            if (sp > 0 && is_minmax(op2)) {   // combined at once: "x MIN", "x NEG MIN"
                if (neg) {
                    nv = make_float4(-nv.x, -nv.y, -nv.z, -nv.w);
                }
                t0 = minmax(t0, nv, (op2 == kMax || op2 == kCMax) ? -1.0f : 1.0f, kblend, op2 >= kCMax ? 0.0f : kgap, grad,
                            &near);
                pc += 1 + neg;
                continue;
            }
            if (sp < kStack) {   // push
                if (sp >= 4) {
                    spill[sp - 4] = t3;
                }
                t3 = t2;
                t2 = t1;
                t1 = t0;
                t0 = nv;
                ++sp;
            }
        }
        if (grad || !near) {
            break;
        }
        grad = 1;
    }
    return sp > 0 ? t0 : sv(-1e30f);
}
}  // namespace sdf

// shape fields: clamp(0.5 + s_l / w) (v2m_phi_f, src/opencl/v2m_seed_body.cl:33-36)
__device__ __noinline__ float shape_phi(const Field* fp, int l, P3 p) {
    const float* prog = fp->gTW;
    return fminf(1.0f, fmaxf(0.0f, 0.5f + sdf::evalv(prog, l, p.x, p.y, p.z, 0).w / prog[1]));
}

// shape fields: (grad psi, psi) of psi_ab = (s_a - s_b) / w (src/opencl/v2m_seed_body.cl:182-221)
__device__ __noinline__ float4 shape_psi(const Field* fp, int a, int b, P3 p, int grad) {
    const float* prog = fp->gTW;
    const float w = prog[1];
    const float4 sa = sdf::evalv(prog, a, p.x, p.y, p.z, grad), sb = sdf::evalv(prog, b, p.x, p.y, p.z, grad);
    return make_float4((sa.x - sb.x) / w, (sa.y - sb.y) / w, (sa.z - sb.z) / w, (sa.w - sb.w) / w);
}

// v2m_phi_f, src/opencl/v2m_seed_body.cl:33-39: v2m_phi_at (src/opencl/v2m_grid_body.cl:
// 189-200), or the shape fields' clamp
__device__ __forceinline__ float phi_point(const Field& f, int l, P3 p) {
    if (f.gm < 0) {
        return shape_phi(f.self, l, p);
    }
    const float u = over_vx(f, p.x), v = over_vy(f, p.y), w = over_vz(f, p.z);
    const int i0 = (int)floorf(u), j0 = (int)floorf(v), k0 = (int)floorf(w);
    float c[8];
    phi_corners(f, l, i0, j0, k0, c);
    return trilerp(c, u - (float)i0, v - (float)j0, w - (float)k0);
}

// the trilinear value and gradient of v2m_psi_grad (src/opencl/v2m_seed_body.cl:
// 233-246) from the corner differences c: (gx, gy, gz, psi)
__device__ __forceinline__ float4 trilerp_grad(const Field& f, const float* c, float fx, float fy, float fz) {
    const float x00 = c[0] + fx * (c[1] - c[0]), x10 = c[2] + fx * (c[3] - c[2]);
    const float x01 = c[4] + fx * (c[5] - c[4]), x11 = c[6] + fx * (c[7] - c[6]);
    const float y0 = x00 + fy * (x10 - x00), y1 = x01 + fy * (x11 - x01);
    const float dx0 = (c[1] - c[0]) + fy * ((c[3] - c[2]) - (c[1] - c[0]));
    const float dx1 = (c[5] - c[4]) + fy * ((c[7] - c[6]) - (c[5] - c[4]));
    float4 g;
    g.x = over_vx(f, dx0 + fz * (dx1 - dx0));
    g.y = over_vy(f, (x10 - x00) + fz * ((x11 - x01) - (x10 - x00)));
    g.z = over_vz(f, y1 - y0);
    g.w = y0 + fz * (y1 - y0);
    return g;
}

// v2m_gray_at, src/opencl/v2m_seed_body.cl:65-93: the trilinear intensity at p,
// (gx, gy, gz, I); a corner outside the grid reads t_0 - W_0. The gradient costs a
// few subtractions more than the value, so both are always formed.
__device__ __forceinline__ float4 gray_at(const Field& f, P3 p) {
    const float u = over_vx(f, p.x), v = over_vy(f, p.y), w = over_vz(f, p.z);
    const int i0 = (int)floorf(u), j0 = (int)floorf(v), k0 = (int)floorf(w);
    const float fx = u - (float)i0, fy = v - (float)j0, fz = w - (float)k0;
    const float out = f.gTW[0] - f.gTW[f.gm];
    float c[8];
    if (i0 >= 0 && j0 >= 0 && k0 >= 0 && i0 + 1 < f.nx && j0 + 1 < f.ny && k0 + 1 < f.nz) {
        const float* b0 = f.gI + vox(f, i0, j0, k0);
        const size_t sy = (size_t)f.nx, sz = (size_t)f.nx * (size_t)f.ny;
        c[0] = b0[0];
        c[1] = b0[1];
        c[2] = b0[sy];
        c[3] = b0[sy + 1];
        c[4] = b0[sz];
        c[5] = b0[sz + 1];
        c[6] = b0[sz + sy];
        c[7] = b0[sz + sy + 1];
    } else {
#pragma unroll
        for (int m = 0; m < 8; ++m) {
            const int i = i0 + (m & 1), j = j0 + ((m >> 1) & 1), k = k0 + (m >> 2);
            c[m] = (i < 0 || j < 0 || k < 0 || i >= f.nx || j >= f.ny || k >= f.nz) ? out : f.gI[vox(f, i, j, k)];
        }
    }
    return trilerp_grad(f, c, fx, fy, fz);
}

// v2m_gray_psi, src/opencl/v2m_seed_body.cl:97-113: for labels a, b of a gray-scale
// volume with |a - b| = 1, psi_ab = +-2 (I - t_k) / W_k (k = min(a, b)) and its
// gradient, (gx, gy, gz, psi). A pair above the last threshold (labels > m, which a
// gray-scale volume does not hold) takes the phi form, so t_k and W_k stay in gTW.
__device__ __forceinline__ bool gray_pair(const Field& f, int a, int b) {
    return f.gm > 0 && (a - b == 1 || b - a == 1) && (a < b ? a : b) < f.gm;
}

__device__ __forceinline__ float4 gray_psi(const Field& f, int a, int b, float4 ig) {
    const int k = a < b ? a : b;
    const float sg = (a > b ? 2.0f : -2.0f) / f.gTW[f.gm + k];
    return make_float4(ig.x * sg, ig.y * sg, ig.z * sg, sg * (ig.w - f.gTW[k]));
}

// v2m_label_of in gray-scale mode (voxmode 0), src/opencl/v2m_seed_body.cl:119-131:
// (the number of thresholds <= I, the runner-up across the nearer threshold in
// membership units)
__device__ __forceinline__ int2 gray_owner(const Field& f, float I) {
    const int m = f.gm;
    int l = 0;
    while (l < m && I >= f.gTW[l]) {
        ++l;
    }
    const float dl = l > 0 ? (I - f.gTW[l - 1]) / f.gTW[m + l - 1] : 1e30f;
    const float du = l < m ? (f.gTW[l] - I) / f.gTW[m + l] : 1e30f;
    return make_int2(l, dl < du ? l - 1 : (l < m ? l + 1 : kNoLab));
}

// v2m_psi, src/opencl/v2m_seed_body.cl:179-205, through a cell cache (the gray-scale
// form for adjacent labels)
__device__ __forceinline__ float psi_at(const Field& f, int a, int b, P3 p, PsiCell& pc) {
    if (f.gm < 0) {
        return shape_psi(f.self, a, b, p, 0).w;
    }
    if (gray_pair(f, a, b)) {
        return gray_psi(f, a, b, gray_at(f, p)).w;
    }
    const float u = over_vx(f, p.x), v = over_vy(f, p.y), w = over_vz(f, p.z);
    const int i0 = (int)floorf(u), j0 = (int)floorf(v), k0 = (int)floorf(w);
    psi_cell_load(f, a, b, i0, j0, k0, pc);
    return trilerp(pc.d, u - (float)i0, v - (float)j0, w - (float)k0);
}

// v2m_psi_grad, src/opencl/v2m_seed_body.cl:211-247: (gx, gy, gz, psi), through a
// cell cache (the gray-scale form for adjacent labels)
__device__ __forceinline__ float4 psi_grad_at(const Field& f, int a, int b, P3 p, PsiCell& pc) {
    if (f.gm < 0) {
        return shape_psi(f.self, a, b, p, 1);
    }
    if (gray_pair(f, a, b)) {
        return gray_psi(f, a, b, gray_at(f, p));
    }
    const float u = over_vx(f, p.x), v = over_vy(f, p.y), w = over_vz(f, p.z);
    const int i0 = (int)floorf(u), j0 = (int)floorf(v), k0 = (int)floorf(w);
    const float fx = u - (float)i0, fy = v - (float)j0, fz = w - (float)k0;
    psi_cell_load(f, a, b, i0, j0, k0, pc);
    return trilerp_grad(f, pc.d, fx, fy, fz);
}

// psi_ab and its gradient at one point (no cache to share)
__device__ __noinline__ float4 psi_grad_point(const Field* fp, int a, int b, P3 p) {
    const Field f = *fp;
    PsiCell pc;
    psi_cell_init(pc);
    return psi_grad_at(f, a, b, p, pc);
}

// The labels of p's brick (v2m_labels_near, src/opencl/v2m_seed_body.cl:42-59) as
// a brick index and a count: bl_lab[b*16 + s] for s < n (b = -1: outside the grid,
// the one label 0).
struct BrickLabels {
    int b, n;
    BrickList list;   // the brick's list (b >= 0)
};

__device__ __forceinline__ BrickLabels brick_labels_at(const Field& f, P3 p) {
    const int i = (int)floorf(over_vx(f, p.x) + 0.5f), j = (int)floorf(over_vy(f, p.y) + 0.5f),
              k = (int)floorf(over_vz(f, p.z) + 0.5f);
    BrickLabels r;
    if (i < 0 || j < 0 || k < 0 || i >= f.nx || j >= f.ny || k >= f.nz) {
        r.b = -1;
        r.n = 1;
        r.list.lo = r.list.hi = make_uint4(0u, 0u, 0u, 0u);
        r.list.cnt = 1;
        r.list.s0 = -1;
        return r;
    }
    r.b = brick_at(f, i, j, k);
    r.list = brick_list(f, r.b);
    r.n = r.list.cnt < kBL ? r.list.cnt : kBL;
    return r;
}

// entry s of a loaded list (a select over its eight words, no memory access)
__device__ __forceinline__ int list_label(const BrickList& bl, int s) {
    const int q = s >> 1;
    const uint4 v = s < 8 ? bl.lo : bl.hi;
    const uint32_t w = (q & 2) ? ((q & 1) ? v.w : v.z) : ((q & 1) ? v.y : v.x);
    return (int)((w >> ((s & 1) * 16)) & 0xFFFFu);
}

__device__ __forceinline__ int brick_label(const Field& f, const BrickLabels& bl, int s) {
    return bl.b < 0 ? 0 : list_label(bl.list, s);
}

// The trilinear cell of a point, and whether it lies in one brick that is the brick
// of the point's rounded voxel (whose labels v2m_labels_near lists): the slot of
// that brick's s-th label is then bl_slot[b] + s, with no search (the labels of a
// brick are distinct, so the search would stop at s).
struct PointCell {
    int i0, j0, k0;
    float fx, fy, fz;
    int s0;     // bl_slot of the shared brick (-1: uniform brick)
    int same;   // the cell lies in the labels' brick
};

__device__ __forceinline__ PointCell point_cell(const Field& f, P3 p, BrickLabels bl) {
    PointCell c;
    const float u = over_vx(f, p.x), v = over_vy(f, p.y), w = over_vz(f, p.z);
    c.i0 = (int)floorf(u);
    c.j0 = (int)floorf(v);
    c.k0 = (int)floorf(w);
    c.fx = u - (float)c.i0;
    c.fy = v - (float)c.j0;
    c.fz = w - (float)c.k0;
    c.same = bl.b >= 0 && cell_in_one_brick(f, c.i0, c.j0, c.k0) && brick_at(f, c.i0, c.j0, c.k0) == bl.b;
    c.s0 = c.same ? bl.list.s0 : -1;
    return c;
}

// the corners of phi_l for the s-th label l of bl at the point's cell (v2m_phi_cell)
__device__ __forceinline__ void label_corners(const Field& f, const PointCell& pc, int s, int l, float* c) {
    if (pc.same) {
        corners_in_brick(f, l, pc.s0 < 0 ? -1 : pc.s0 + s, 1, pc.i0, pc.j0, pc.k0, c);
    } else {
        phi_corners(f, l, pc.i0, pc.j0, pc.k0, c);
    }
}

// phi of the s-th label l of bl at the point (v2m_phi_at)
__device__ __forceinline__ float label_phi(const Field& f, const PointCell& pc, int s, int l) {
    float c[8];
    label_corners(f, pc, s, l, c);
    return trilerp(c, pc.fx, pc.fy, pc.fz);
}

// v2m_label_of (voxmode 0), src/opencl/v2m_seed_body.cl:116-175: (best, second)
// shape fields: the brick's labels ranked by s itself (src/opencl/v2m_seed_body.cl:137-160)
__device__ __noinline__ int2 shape_owner(const Field* fp, P3 p) {
    const Field f = *fp;
    const BrickLabels bl = brick_labels_at(f, p);
    int best = brick_label(f, bl, 0), sec = kNoLab;
    if (bl.n == 1) {
        return make_int2(best, kNoLab);
    }
    float pb = -1e30f, ps = pb;
    for (int s = 0; s < bl.n; ++s) {
        const int l = brick_label(f, bl, s);
        const float v = sdf::evalv(f.gTW, l, p.x, p.y, p.z, 0).w;
        if (v > pb) {
            ps = pb;
            sec = best;
            pb = v;
            best = l;
        } else if (v > ps) {
            ps = v;
            sec = l;
        }
    }
    return make_int2(best, sec);
}

__device__ __forceinline__ int2 owner_at(const Field& f, P3 p) {
    if (f.gm > 0) {
        return gray_owner(f, gray_at(f, p).w);
    }
    if (f.gm < 0) {
        return shape_owner(f.self, p);
    }
    const BrickLabels bl = brick_labels_at(f, p);
    int best = brick_label(f, bl, 0), sec = kNoLab;
    float pb = -1.0f, ps = pb;
    if (bl.n == 1) {
        return make_int2(best, kNoLab);
    }
    const PointCell pc = point_cell(f, p, bl);
    for (int s = 0; s < bl.n; ++s) {
        const int l = brick_label(f, bl, s);
        const float v = label_phi(f, pc, s, l);
        if (v > pb) {
            ps = pb;
            sec = best;
            pb = v;
            best = l;
        } else if (v > ps) {
            ps = v;
            sec = l;
        }
    }
    return make_int2(best, sec);
}

// owner_at, and when the owner is label a with a runner-up, psi_{a,second} and its
// gradient at p from the same corners (v2m_psi_grad, src/opencl/v2m_seed_body.cl:
// 211-247): the interior branch of the trapped step needs exactly these.
struct OwnerPsi {
    int best, sec;
    float4 g;   // (grad psi, psi) when best == a and sec != kNoLab
};

__device__ __forceinline__ OwnerPsi owner_psi_at(const Field& f, int a, P3 p) {
    OwnerPsi r;
    if (f.gm > 0) {   // the runner-up is adjacent: psi from the same intensity
        const float4 ig = gray_at(f, p);
        const int2 o = gray_owner(f, ig.w);
        r.best = o.x;
        r.sec = o.y;
        r.g = (o.x == a && o.y != kNoLab) ? gray_psi(f, a, o.y, ig) : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        return r;
    }
    if (f.gm < 0) {   // the owner by the values, then psi of (a, second) with the gradients
        const int2 o = shape_owner(f.self, p);
        r.best = o.x;
        r.sec = o.y;
        r.g = (o.x == a && o.y != kNoLab) ? shape_psi(f.self, a, o.y, p, 1) : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        return r;
    }
    const BrickLabels bl = brick_labels_at(f, p);
    r.best = brick_label(f, bl, 0);
    r.sec = kNoLab;
    r.g = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (bl.n == 1) {
        return r;
    }
    const PointCell pc = point_cell(f, p, bl);
    float pb = -1.0f, ps = pb;
    float cb[8], cs[8];
#pragma unroll
    for (int m = 0; m < 8; ++m) {
        cb[m] = cs[m] = 0.0f;
    }
    for (int s = 0; s < bl.n; ++s) {
        const int l = brick_label(f, bl, s);
        float c[8];
        label_corners(f, pc, s, l, c);
        const float v = trilerp(c, pc.fx, pc.fy, pc.fz);
        if (v > pb) {
            ps = pb;
            r.sec = r.best;
            pb = v;
            r.best = l;
#pragma unroll
            for (int m = 0; m < 8; ++m) {
                cs[m] = cb[m];
                cb[m] = c[m];
            }
        } else if (v > ps) {
            ps = v;
            r.sec = l;
#pragma unroll
            for (int m = 0; m < 8; ++m) {
                cs[m] = c[m];
            }
        }
    }
    if (r.best == a && r.sec != kNoLab) {
        // from the second label on, cs holds the runner-up's corners (label 0 becomes
        // best at s = 0; at s = 1 either it moves to second or label 1 takes second)
        float d[8];
#pragma unroll
        for (int m = 0; m < 8; ++m) {
            d[m] = cb[m] - cs[m];
        }
        r.g = trilerp_grad(f, d, pc.fx, pc.fy, pc.fz);
    }
    return r;
}

__device__ __forceinline__ P3 along(P3 p, float sg, float t, P3 u) {
    return p3(p.x + sg * t * u.x, p.y + sg * t * u.y, p.z + sg * t * u.z);
}

// v2m_bisect_dir, src/opencl/v2m_seed_body.cl:255-320 (doubling probes, then Illinois).
// On success the returned point is the last point probed, so pc then holds its cell.
__device__ __forceinline__ P3ok bracket_along(const Field& f, int a, int b, P3 p, P3 u, float v0, float rad, float tg,
                                              PsiCell& pc) {
    P3ok r;
    r.p = p;
    r.ok = 1;
    if (v0 == 0.0f) {
        return r;
    }
    const float sg = v0 > 0.0f ? -1.0f : 1.0f;
    float t0 = 0.0f, t1 = -1.0f, f0 = v0, f1 = 0.0f;
    float t = (tg > 0.0f && tg < rad) ? tg : 0.25f * rad;
    for (int k = 0; k < 8; ++k) {
        const float v = psi_at(f, a, b, along(p, sg, t, u), pc);
        if ((v > 0.0f) != (v0 > 0.0f) || v == 0.0f) {
            t1 = t;
            f1 = v;
            break;
        }
        t0 = t;
        f0 = v;
        if (t >= rad) {
            break;
        }
        t = fminf(2.0f * t, rad);
    }
    if (t1 < 0.0f) {
        r.ok = 0;
        return r;
    }
    int side = 0;
    for (int it = 0; it < 8 && f1 != 0.0f && fabsf(t1 - t0) > 1e-4f * rad; ++it) {
        const float tm = (f0 != f1) ? t1 - f1 * (t1 - t0) / (f1 - f0) : 0.5f * (t0 + t1);
        const float v = psi_at(f, a, b, along(p, sg, tm, u), pc);
        if ((v > 0.0f) == (f1 > 0.0f)) {
            t1 = tm;
            f1 = v;
            if (side == -1) {
                f0 *= 0.5f;
            }
            side = -1;
        } else {
            t0 = t1;
            f0 = f1;
            t1 = tm;
            f1 = v;
            side = 1;
        }
    }
    r.p.x = p.x + sg * t1 * u.x;
    r.p.y = p.y + sg * t1 * u.y;
    r.p.z = p.z + sg * t1 * u.z;
    return r;
}

// v2m_project1, src/opencl/v2m_seed_body.cl:324-335 (ok = 0: p unchanged), with the
// a|b cell cache of the caller
__device__ __forceinline__ P3ok onto_interface_at(const Field& f, int a, int b, P3 p, float rad, PsiCell& pc) {
    const float4 g = psi_grad_at(f, a, b, p, pc);
    const float gn = sqrtf(g.x * g.x + g.y * g.y + g.z * g.z);
    if (gn < 1e-9f) {
        P3ok r;
        r.p = p;
        r.ok = 0;
        return r;
    }
    const P3 u = p3(g.x / gn, g.y / gn, g.z / gn);
    return bracket_along(f, a, b, p, u, g.w, rad, 1.5f * fabsf(g.w) / gn, pc);
}

__device__ __noinline__ P3ok onto_interface(const Field* fp, int a, int b, P3 p, float rad) {
    const Field f = *fp;
    PsiCell pc;
    psi_cell_init(pc);
    return onto_interface_at(f, a, b, p, rad, pc);
}

// v2m_project2 (gm == 0), src/opencl/v2m_seed_body.cl:342-420 (ok = 0: p unchanged)
__device__ __noinline__ P3ok onto_junction(const Field* fp, int a, int b, int c, P3 p, float rad) {
    const Field f = *fp;
    P3ok fail;
    fail.p = p;
    fail.ok = 0;
    P3 q = p;
    PsiCell pab, pac;   // the a|b and a|c cells: each projection ends where its cache points
    psi_cell_init(pab);
    psi_cell_init(pac);
    if (f.gm < 0) {
        // shape fields: of the three pairs' constraints, the two with the most
        // independent gradients, their shared label first (src/opencl/v2m_seed_body.cl:345-375)
        const float4 gab = shape_psi(fp, a, b, q, 1), gac = shape_psi(fp, a, c, q, 1), gbc = shape_psi(fp, b, c, q, 1);
        const float xa[3] = { gab.y * gac.z - gab.z * gac.y, gab.z * gac.x - gab.x * gac.z, gab.x * gac.y - gab.y * gac.x };
        const float xb[3] = { gab.y * gbc.z - gab.z * gbc.y, gab.z * gbc.x - gab.x * gbc.z, gab.x * gbc.y - gab.y * gbc.x };
        const float xc[3] = { gac.y * gbc.z - gac.z * gbc.y, gac.z * gbc.x - gac.x * gbc.z, gac.x * gbc.y - gac.y * gbc.x };
        const float na = xa[0] * xa[0] + xa[1] * xa[1] + xa[2] * xa[2], nb = xb[0] * xb[0] + xb[1] * xb[1] + xb[2] * xb[2],
                    nc = xc[0] * xc[0] + xc[1] * xc[1] + xc[2] * xc[2];
        if (nb > na && nb >= nc) {   // pivot b
            const int t = a;
            a = b;
            b = t;
        } else if (nc > na && nc > nb) {   // pivot c
            const int t = a;
            a = c;
            c = b;
            b = t;
        }
    }
    for (int it = 0; it < 3; ++it) {
        const P3ok r1 = onto_interface_at(f, a, b, q, rad, pab);
        if (!r1.ok) {
            return fail;
        }
        q = r1.p;
        const float4 g1 = psi_grad_at(f, a, b, q, pab);
        const float4 g2 = psi_grad_at(f, a, c, q, pac);
        const float v2 = g2.w;
        const float a11 = g1.x * g1.x + g1.y * g1.y + g1.z * g1.z;
        if (a11 < 1e-18f) {
            return fail;
        }
        const float pr = (g1.x * g2.x + g1.y * g2.y + g1.z * g2.z) / a11;
        P3 u = p3(g2.x - pr * g1.x, g2.y - pr * g1.y, g2.z - pr * g1.z);
        const float un = sqrtf(u.x * u.x + u.y * u.y + u.z * u.z);
        if (un < 1e-9f) {
            return fail;
        }
        u.x /= un;
        u.y /= un;
        u.z /= un;
        const P3ok r2 = bracket_along(f, a, c, q, u, v2, rad, 1.5f * fabsf(v2) / un, pac);
        if (!r2.ok) {
            return fail;
        }
        q = r2.p;
    }
    const P3ok r3 = onto_interface_at(f, a, b, q, rad, pab);
    if (!r3.ok) {
        return fail;
    }
    q = r3.p;
    const float m2 = (q.x - p.x) * (q.x - p.x) + (q.y - p.y) * (q.y - p.y) + (q.z - p.z) * (q.z - p.z);
    if (m2 > rad * rad) {
        return fail;
    }
    P3ok r;
    r.p = q;
    r.ok = 1;
    return r;
}

// ---- hash ------------------------------------------------------------------------------

// v2m_h_at, src/opencl/v2m_particle_body.cl:30-36
__device__ __forceinline__ float size_at(const Field& f, float px, float py, float pz) {
    int i = (int)floorf(over_vx(f, px) + 0.5f), j = (int)floorf(over_vy(f, py) + 0.5f),
        k = (int)floorf(over_vz(f, pz) + 0.5f);
    i = i < 0 ? 0 : (i >= f.nx ? f.nx - 1 : i);
    j = j < 0 ? 0 : (j >= f.ny ? f.ny - 1 : j);
    k = k < 0 ? 0 : (k >= f.nz ? f.nz - 1 : k);
    return f.h[vox(f, i, j, k)];
}

// v2m_level_of, src/opencl/v2m_particle_body.cl:38-46
__device__ __forceinline__ int level_for(const Hash& H, float R) {
    int L = 0;
    while (L + 1 < H.nlev && H.b0 * (float)(1 << L) < R) {
        ++L;
    }
    return L;
}

// v2m_bin_coord, src/opencl/v2m_particle_body.cl:48-52
__device__ __forceinline__ int cell_coord(const Hash& H, int L, float x, float o, int ax) {
    const int c = (int)floorf((x - o) / (H.b0 * (float)(1 << L)));
    const int n = H.dim[L][ax];
    return c < 0 ? 0 : (c >= n ? n - 1 : c);
}

// v2m_bin_key, src/opencl/v2m_particle_body.cl:54-58
__device__ __forceinline__ int cell_key(const Hash& H, int L, float px, float py, float pz) {
    const int cx = cell_coord(H, L, px, H.ox, 0), cy = cell_coord(H, L, py, H.oy, 1), cz = cell_coord(H, L, pz, H.oz, 2);
    return H.off[L] + cx + H.dim[L][0] * (cy + H.dim[L][1] * cz);
}

// ---- truss -----------------------------------------------------------------------------

// The inputs of one K-nearest search (v2m_neighbors, src/opencl/v2m_particle_body.cl:73-151).
// Nodes sit in slots renumbered in cell order at each rebuild.
struct TrussIn {
    const float4* __restrict__ X;        // current positions + h, by slot
    const uint16_t* __restrict__ lab;    // by slot
    const uint8_t* __restrict__ typ;     // current types, by slot
    const int32_t* __restrict__ cstart;  // cell starts (slots) as of the last rebuild
    const float4* __restrict__ Ps;       // positions + h as of the last rebuild, by slot
    const int2* __restrict__ Ms;         // (id, typ | lab << 8) as of the last rebuild, by slot
    const int32_t* __restrict__ slot_of; // id -> slot
    const int32_t* __restrict__ order;   // the slots in the K-nearest search's query order
    float t, skin;
};

// v2m_force, src/opencl/v2m_particle_body.cl:161-184: the repulsive bars of node i,
// summed in list order (the fixed order that makes the sum reproducible)
__device__ __forceinline__ float4 bar_force(const float4* __restrict__ X, const uint8_t* __restrict__ typ,
                                            const int32_t* __restrict__ nbrT, const int32_t* __restrict__ nnb,
                                            float fscale, float fsurf, int i, int n_nodes) {
    const float4 xi = X[i];
    const float px = xi.x, py = xi.y, pz = xi.z, hi = xi.w;
    const int ti = typ[i];
    float F0 = 0.0f, F1 = 0.0f, F2 = 0.0f, F3 = 0.0f;
    const int m = nnb[i];
    // the whole list read at once, then eight bars at a time: their loads are issued
    // together, the sum keeps list order (entries past the list read the node itself and
    // are not summed)
    constexpr int kB = 8;
    static_assert(TN_K % kB == 0, "the list length is a multiple of the batch");
    int jl[TN_K];
#pragma unroll
    for (int u = 0; u < TN_K; ++u) {
        jl[u] = u < m ? nbrT[(size_t)u * n_nodes + i] : i;
    }
#pragma unroll
    for (int k0 = 0; k0 < TN_K; k0 += kB) {
        if (k0 >= m) {
            break;
        }
        float4 xj[kB];
        int tj[kB];
#pragma unroll
        for (int u = 0; u < kB; ++u) {
            xj[u] = X[jl[k0 + u]];
            tj[u] = typ[jl[k0 + u]];
        }
#pragma unroll
        for (int u = 0; u < kB; ++u) {
            if (k0 + u < m) {
                const float ex = px - xj[u].x, ey = py - xj[u].y, ez = pz - xj[u].z;
                const float l = sqrtf(ex * ex + ey * ey + ez * ez);
                const float l0 = ((ti != TN_INTERIOR && tj[u] != TN_INTERIOR) ? fsurf : fscale) * 0.5f * (hi + xj[u].w);
                if (!(l >= l0 || l <= 0.0f)) {
                    const float fr = (l0 - l) / l;
                    F0 += fr * ex;
                    F1 += fr * ey;
                    F2 += fr * ez;
                    F3 += 1.0f;
                }
            }
        }
    }
    return make_float4(F0, F1, F2, F3);
}

// ---- the trapped move --------------------------------------------------------------------

// v2m_crossing, src/opencl/v2m_particle_body.cl:188-212
__device__ __forceinline__ P3 segment_crossing_at(const Field& f, int a, int b, P3 p, P3 q, PsiCell& pc) {
    float t0 = 0.0f, t1 = 1.0f;
    float v0 = psi_at(f, a, b, p, pc);
    float v1 = psi_at(f, a, b, q, pc);
    for (int it = 0; it < 6; ++it) {
        const float tm = 0.5f * (t0 + t1);
        const float vm = psi_at(f, a, b, p3(p.x + tm * (q.x - p.x), p.y + tm * (q.y - p.y), p.z + tm * (q.z - p.z)), pc);
        if (vm > 0.0f) {
            t0 = tm;
            v0 = vm;
        } else {
            t1 = tm;
            v1 = vm;
        }
    }
    const float t = (v0 - v1) != 0.0f ? t0 + (t1 - t0) * v0 / (v0 - v1) : 0.5f * (t0 + t1);
    return p3(p.x + t * (q.x - p.x), p.y + t * (q.y - p.y), p.z + t * (q.z - p.z));
}

// v2m_tangential, src/opencl/v2m_particle_body.cl:215-229: v less its part along n_ab at p
__device__ __forceinline__ P3 drop_normal_at(const Field& f, int a, int b, P3 p, P3 v, PsiCell& pc) {
    const float4 g = psi_grad_at(f, a, b, p, pc);
    const float gn2 = g.x * g.x + g.y * g.y + g.z * g.z;
    if (gn2 < 1e-12f) {
        return v;
    }
    const float vn = (v.x * g.x + v.y * g.y + v.z * g.z) / gn2;
    return p3(v.x - vn * g.x, v.y - vn * g.y, v.z - vn * g.z);
}

// v2m_third_label, src/opencl/v2m_particle_body.cl:232-260. The brick test runs
// inline; the field comparisons (a third label is near: rare) are a call.
__device__ __noinline__ int competing_label_near(const Field* fp, int a, int b, P3 q, BrickLabels bl) {
    const Field f = *fp;
    const float pa = phi_point(f, a, q);
    int c = kNoLab;
    float pc = 0.5f * pa;
    const PointCell cell = point_cell(f, q, bl);
    for (int s = 0; s < bl.n; ++s) {
        const int l = brick_label(f, bl, s);
        if (l != a && l != b) {
            const float w = f.gm < 0 ? phi_point(f, l, q) : label_phi(f, cell, s, l);
            if (w > pc) {
                pc = w;
                c = l;
            }
        }
    }
    return c;
}

__device__ __forceinline__ int competing_label(const Field& f, int a, int b, P3 q) {
    const BrickLabels bl = brick_labels_at(f, q);
    int other = 0;
    for (int s = 0; s < bl.n; ++s) {
        const int l = brick_label(f, bl, s);
        other |= l != a && l != b;
    }
    if (!other) {
        return kNoLab;
    }
    return competing_label_near(f.self, a, b, q, bl);
}

// v2m_valid_on, src/opencl/v2m_particle_body.cl:266-291 (the same split)
__device__ __noinline__ int labels_dominate_near(const Field* fp, int a, int b, int c, P3 q, BrickLabels bl) {
    const Field f = *fp;
    float mine = fminf(phi_point(f, a, q), phi_point(f, b, q));
    if (c != kNoLab) {
        mine = fminf(mine, phi_point(f, c, q));
    }
    const PointCell cell = point_cell(f, q, bl);
    for (int s = 0; s < bl.n; ++s) {
        const int l = brick_label(f, bl, s);
        if (l != a && l != b && l != c && (f.gm < 0 ? phi_point(f, l, q) : label_phi(f, cell, s, l)) > mine + 0.02f) {
            return 0;
        }
    }
    return 1;
}

__device__ __forceinline__ int labels_dominate(const Field& f, int a, int b, int c, P3 q) {
    const BrickLabels bl = brick_labels_at(f, q);
    int other = 0;
    for (int s = 0; s < bl.n; ++s) {
        const int l = brick_label(f, bl, s);
        other |= l != a && l != b && l != c;
    }
    if (!other) {
        return 1;
    }
    return labels_dominate_near(f.self, a, b, c, q, bl);
}

constexpr float kDeferred = -1.0f;   // trapped_step<true>: the node needs the general path

__device__ __forceinline__ float dist2(P3 q, P3 p) {
    return (q.x - p.x) * (q.x - p.x) + (q.y - p.y) * (q.y - p.y) + (q.z - p.z) * (q.z - p.z);
}

// A node's mutable state during one move (registers; written back by the caller).
struct NodeState {
    P3 p;           // position in, new position out
    float hn;       // h at the position
    int ty;         // type: typ[i] as stored
    int b, c;       // partners as stored: part[2i], part[2i+1]
};

// v2m_move_step (smooth mode), src/opencl/v2m_particle_body.cl:659-833: apply the
// capped step s with interface trapping. Returns |dp| / h (0: the node stayed;
// its type and partners may still have changed, as in the reference).
// The general tail below for an interior or interface node, without its out-of-line
// calls: one pass over the labels of q's brick evaluates both v2m_third_label
// (competing_label_near) and v2m_valid_on (labels_dominate_near) at q, with the same
// values. A third label found (the node would try to join a junction) returns
// kDeferred. Otherwise the step fails if it is too long or another label dominates
// at q (the second third-label search then finds none either: same labels, same
// point), and ends as the general tail ends if not.
__device__ __forceinline__ float lean_tail(const Field& f, float h, float maxstep, float snap, int a, int typ0,
                                           int ty, int b, P3 p, P3 q, NodeState& ns) {
    bool dominated = false;
    if (ty == TN_INTERFACE) {
        const BrickLabels bl = brick_labels_at(f, q);
        int other = 0;
        for (int s = 0; s < bl.n; ++s) {
            const int l = brick_label(f, bl, s);
            other |= l != a && l != b;
        }
        if (other) {
            const PointCell cell = point_cell(f, q, bl);
            const float pa = phi_point(f, a, q), pb = phi_point(f, b, q);
            const float mine = fminf(pa, pb);
            float pc = 0.5f * pa;
            int cc = kNoLab;
            for (int s = 0; s < bl.n; ++s) {
                const int l = brick_label(f, bl, s);
                if (l != a && l != b) {
                    const float w = f.gm < 0 ? phi_point(f, l, q) : label_phi(f, cell, s, l);
                    if (w > pc) {
                        pc = w;
                        cc = l;
                    }
                    dominated = dominated || w > mine + 0.02f;
                }
            }
            if (cc != kNoLab) {
                return kDeferred;
            }
        }
    }
    const float lim = (maxstep + fmaxf(snap, 0.5f)) * h;
    const float m2 = dist2(q, p);
    if (!(m2 <= lim * lim)) {
        if (typ0 == TN_INTERIOR) {
            ns.b = kNoLab;
        }
        return 0.0f;
    }
    if (ty != TN_INTERIOR && dominated) {
        ns.ty = typ0 == TN_INTERIOR ? TN_INTERIOR : ty;
        if (ns.ty == TN_INTERIOR) {
            ns.b = kNoLab;
        }
        return 0.0f;
    }
    ns.p = q;
    ns.ty = ty;
    ns.hn = size_at(f, q.x, q.y, q.z);
    return sqrtf(m2) / h;
}

// Step modes: the general path, or a lean one for nodes known to be interface
// (kLeanInterface) or interior (kLeanInterior) nodes: the other branches are compiled
// out, and so is the rare path (a third label strong enough for a junction where the
// node lands), for which the step returns kDeferred and leaves the stored state
// untouched; the caller hands such a node to the general path.
constexpr int kGeneral = 0, kLeanInterface = 1, kLeanInterior = 2;

template <int kMode>
__device__ __forceinline__ float trapped_step(const Field& f, float h, P3 s, float maxstep, float snap, int a,
                                              NodeState& ns) {
    const P3 p = ns.p;
    P3 q;
    // typ[i] as stored until the final write
    const int typ0 = kMode == kLeanInterface ? TN_INTERFACE : (kMode == kLeanInterior ? TN_INTERIOR : ns.ty);
    int ty = typ0;
    int b = ns.b;             // the locals b, c of the reference; ns.b / ns.c are the stored part[]
    const int c = ns.c;

    if (kMode != kLeanInterface && ty == TN_INTERIOR) {
        q = p3(p.x + s.x, p.y + s.y, p.z + s.z);
        const int2 own = owner_at(f, q);
        const int l = own.x, sec = own.y;
        if (l != a) {
            b = l;
            PsiCell pc;   // crossing, glide and projection stay near the crossing point
            psi_cell_init(pc);
            const P3 x = segment_crossing_at(f, a, b, p, q, pc);
            const P3 r = drop_normal_at(f, a, b, x, p3(q.x - x.x, q.y - x.y, q.z - x.z), pc);
            q = p3(x.x + r.x, x.y + r.y, x.z + r.z);
            const P3ok pr = onto_interface_at(f, a, b, q, 0.5f * h, pc);
            q = pr.ok ? pr.p : x;
            ty = TN_INTERFACE;
            ns.b = b;
        } else if (sec != kNoLab && snap > 0.0f) {
            PsiCell pc;
            psi_cell_init(pc);
            const float4 gq = psi_grad_at(f, a, sec, q, pc);
            const float v = gq.w;
            const float gn = sqrtf(gq.x * gq.x + gq.y * gq.y + gq.z * gq.z);
            if (gn > 1e-9f && v / gn < snap * h) {
                const P3ok pr = onto_interface_at(f, a, sec, q, 1.2f * snap * h, pc);
                if (pr.ok) {
                    q = pr.p;
                    ty = TN_INTERFACE;
                    b = sec;
                    ns.b = sec;
                }
            }
        }
    } else if (kMode == kLeanInterface || ty == TN_INTERFACE) {
        PsiCell pc;   // the glide and its projection probe the same few cells
        psi_cell_init(pc);
        s = drop_normal_at(f, a, b, p, s, pc);
        q = p3(p.x + s.x, p.y + s.y, p.z + s.z);
        const P3ok pr = onto_interface_at(f, a, b, q, 0.5f * h, pc);
        if (!pr.ok) {
            return 0.0f;
        }
        q = pr.p;
    } else {
        const float4 g1 = psi_grad_point(f.self, a, b, p);
        const float4 g2 = psi_grad_point(f.self, a, c, p);
        const float t0 = g1.y * g2.z - g1.z * g2.y;
        const float t1 = g1.z * g2.x - g1.x * g2.z;
        const float t2 = g1.x * g2.y - g1.y * g2.x;
        const float tl2 = t0 * t0 + t1 * t1 + t2 * t2;
        const float st = tl2 > 1e-20f ? (s.x * t0 + s.y * t1 + s.z * t2) / tl2 : 0.0f;
        q = p3(p.x + st * t0, p.y + st * t1, p.z + st * t2);
        const P3ok pr = onto_junction(f.self, a, b, c, q, 0.5f * h);
        if (!pr.ok) {
            return 0.0f;
        }
        q = pr.p;
    }

    if (kMode != kGeneral) {
        return lean_tail(f, h, maxstep, snap, a, typ0, ty, b, p, q, ns);
    }

    if (ty == TN_INTERFACE) {
        const int cc = competing_label(f, a, b, q);
        if (cc != kNoLab) {
            const P3ok pr = onto_junction(f.self, a, b, cc, q, 0.5f * h);
            if (pr.ok) {
                q = pr.p;
                ty = TN_JUNCTION;
                ns.c = cc;
            }
        }
    }

    {
        const float lim = (maxstep + fmaxf(snap, 0.5f)) * h;
        const float m2 = dist2(q, p);
        if (!(m2 <= lim * lim)) {
            if (typ0 == TN_INTERIOR) {
                ns.b = kNoLab;
            }
            return 0.0f;
        }
    }

    if (ty != TN_INTERIOR && !labels_dominate(f, a, ns.b, ty >= TN_JUNCTION ? ns.c : kNoLab, q)) {
        if (ty == TN_INTERFACE || typ0 == TN_INTERIOR) {
            const int cc = competing_label(f, a, ns.b, q);
            int joined = 0;
            if (cc != kNoLab) {
                const P3ok pr = onto_junction(f.self, a, ns.b, cc, q, 0.5f * h);
                if (pr.ok && labels_dominate(f, a, ns.b, cc, pr.p)) {
                    q = pr.p;
                    joined = 1;
                }
            }
            if (joined) {
                ty = TN_JUNCTION;
                ns.c = cc;
            } else {
                ns.ty = typ0 == TN_INTERIOR ? TN_INTERIOR : ty;
                if (ns.ty == TN_INTERIOR) {
                    ns.b = kNoLab;
                }
                return 0.0f;
            }
        } else {
            return 0.0f;
        }
    }

    const float mv = sqrtf(dist2(q, p));
    ns.p = q;
    ns.ty = ty;
    ns.hn = size_at(f, q.x, q.y, q.z);
    return mv / h;
}

// ---- the voxel trap: the exact staircase of the label volume ---------------------------
// (v2m_move_voxel and its helpers, src/opencl/v2m_particle_body.cl:373-650: interior
// nodes walk their step voxel by voxel and stop on the first face into another label;
// interface and junction nodes take the nearest point on their a|b faces or a|b|c
// edges. Integer label reads only; no phi.)

// v2m_label_at, src/opencl/v2m_grid_body.cl:72-78 (0 outside the grid)
__device__ __forceinline__ int label_or_zero(const Field& f, int i, int j, int k) {
    if (i < 0 || j < 0 || k < 0 || i >= f.nx || j >= f.ny || k >= f.nz) {
        return 0;
    }
    return f.L[vox(f, i, j, k)];
}

__device__ __forceinline__ float over_v(const Field& f, float x, int ax) {
    return ax == 0 ? over_vx(f, x) : (ax == 1 ? over_vy(f, x) : over_vz(f, x));
}

// v2m_dda_hit, src/opencl/v2m_particle_body.cl:373-431
__device__ float dda_hit(const Field& f, int a, const float* p, const float* s, float* x, int* b) {
    const float vs[3] = { f.vx, f.vy, f.vz };
    float u[3], w[3];
    int id[3];
    for (int k = 0; k < 3; ++k) {
        u[k] = over_v(f, p[k], k) + 0.5f;
        w[k] = over_v(f, s[k], k);
        id[k] = (int)floorf(u[k]);
    }
    int l = label_or_zero(f, id[0], id[1], id[2]);
    if (l != a) {
        x[0] = p[0];
        x[1] = p[1];
        x[2] = p[2];
        *b = l;
        return 0.0f;
    }
    float t = 0.0f;
    for (int it = 0; it < 64; ++it) {
        float ht[3];
        for (int k = 0; k < 3; ++k) {
            ht[k] = w[k] != 0.0f ? fabsf(((float)id[k] + (w[k] > 0.0f ? 1.0f : 0.0f) - u[k]) / w[k]) : 1e30f;
        }
        const float dt = fminf(fminf(ht[0], ht[1]), ht[2]);
        const int ax = dt == ht[0] ? 0 : (dt == ht[1] ? 1 : 2);
        if (t + dt >= 1.0f) {
            return -1.0f;
        }
        t += dt;
        for (int k = 0; k < 3; ++k) {
            u[k] += dt * w[k];
        }
        u[ax] = (float)id[ax] + (w[ax] > 0.0f ? 1.0f : 0.0f);
        id[ax] += w[ax] > 0.0f ? 1 : -1;
        l = label_or_zero(f, id[0], id[1], id[2]);
        if (l != a) {
            for (int k = 0; k < 3; ++k) {
                x[k] = (u[k] - 0.5f) * vs[k];
            }
            *b = l;
            return t;
        }
    }
    return -1.0f;
}

// v2m_vox_nearest, src/opencl/v2m_particle_body.cl:437-529
__device__ float vox_nearest(const Field& f, int a, int b, int c, const float* q, int R, float* out) {
    const float vs[3] = { f.vx, f.vy, f.vz };
    float u[3];
    int ci[3];
    for (int k = 0; k < 3; ++k) {
        u[k] = over_v(f, q[k], k) + 0.5f;
        ci[k] = (int)floorf(u[k]);
    }
    float best = -1.0f;
    for (int vk = ci[2] - R; vk <= ci[2] + R; ++vk) {
        for (int vj = ci[1] - R; vj <= ci[1] + R; ++vj) {
            for (int vi = ci[0] - R; vi <= ci[0] + R; ++vi) {
                if (label_or_zero(f, vi, vj, vk) != a) {
                    continue;
                }
                const int v[3] = { vi, vj, vk };
                for (int ax = 0; ax < 3; ++ax) {
                    const int o1 = (ax + 1) % 3, o2 = (ax + 2) % 3;
                    if (c == kNoLab) {   // the two faces normal to ax
                        for (int sd = 0; sd < 2; ++sd) {
                            int nn[3] = { v[0], v[1], v[2] };
                            nn[ax] += sd ? 1 : -1;
                            if (label_or_zero(f, nn[0], nn[1], nn[2]) != b) {
                                continue;
                            }
                            float y[3];
                            y[ax] = (float)(v[ax] + sd);
                            y[o1] = fminf(fmaxf(u[o1], (float)v[o1]), (float)(v[o1] + 1));
                            y[o2] = fminf(fmaxf(u[o2], (float)v[o2]), (float)(v[o2] + 1));
                            float d2 = 0.0f;
                            for (int k = 0; k < 3; ++k) {
                                y[k] = (y[k] - 0.5f) * vs[k];
                                d2 += (y[k] - q[k]) * (y[k] - q[k]);
                            }
                            if (best < 0.0f || d2 < best) {
                                best = d2;
                                out[0] = y[0];
                                out[1] = y[1];
                                out[2] = y[2];
                            }
                        }
                    } else {   // the four edges parallel to ax
                        for (int e = 0; e < 4; ++e) {
                            const int e1 = e & 1, e2 = e >> 1;
                            int hb = 0, hc = 0;
                            for (int m = 0; m < 4; ++m) {
                                int nn[3] = { v[0], v[1], v[2] };
                                nn[o1] += e1 - 1 + (m & 1);
                                nn[o2] += e2 - 1 + (m >> 1);
                                const int ln = label_or_zero(f, nn[0], nn[1], nn[2]);
                                hb |= ln == b;
                                hc |= ln == c;
                            }
                            if (!hb || !hc) {
                                continue;
                            }
                            float y[3];
                            y[ax] = fminf(fmaxf(u[ax], (float)v[ax]), (float)(v[ax] + 1));
                            y[o1] = (float)(v[o1] + e1);
                            y[o2] = (float)(v[o2] + e2);
                            float d2 = 0.0f;
                            for (int k = 0; k < 3; ++k) {
                                y[k] = (y[k] - 0.5f) * vs[k];
                                d2 += (y[k] - q[k]) * (y[k] - q[k]);
                            }
                            if (best < 0.0f || d2 < best) {
                                best = d2;
                                out[0] = y[0];
                                out[1] = y[1];
                                out[2] = y[2];
                            }
                        }
                    }
                }
            }
        }
    }
    return best;
}

constexpr int kVoxBoxR = 4;   // scans up to this reach read their labels from a box in LDS
constexpr int kVoxBox = (2 * kVoxBoxR + 3) * (2 * kVoxBoxR + 3) * (2 * kVoxBoxR + 3);

// vox_nearest with the label reads of its (2R+1)^3 voxels split over a wave (every
// lane active, the same arguments in every lane). Each lane tests one voxel of a
// group of 64 and marks which of its faces (or edges) qualify. The serial scan
// measures each candidate from q, and, when out is q itself (as the callers that move
// the node pass it), from the last point it accepted: then the wave walks the marked
// ones in the serial scan's order, every lane alike. Otherwise each lane keeps its
// first best and the wave takes the smallest (d^2, scan position). Either way the
// serial choice is kept, and the result is the same in every lane.

__device__ float vox_nearest_wave(const Field& f, int a, int b, int c, const float* q, int R, float* out) {
    __shared__ uint16_t box[4][kVoxBox];   // a wave's labels around its scan (4 waves a block)
    const int lane = threadIdx.x & 63;
    const bool chained = out == q;
    const float vs[3] = { f.vx, f.vy, f.vz };
    float u[3], qc[3];
    int ci[3];
    for (int k = 0; k < 3; ++k) {
        u[k] = over_v(f, q[k], k) + 0.5f;
        ci[k] = (int)floorf(u[k]);
        qc[k] = q[k];
    }
    const int W = 2 * R + 1, total = W * W * W, nsub = c == kNoLab ? 2 : 4;
    // The labels of the scan's voxels and their neighbours, (W + 2)^3 of them, read at
    // once (independent loads) into LDS; reaches beyond the box read them directly.
    const bool boxed = R <= kVoxBoxR;
    const int B = W + 2, x0 = ci[0] - R - 1, y0 = ci[1] - R - 1, z0 = ci[2] - R - 1;
    uint16_t* bx = box[(threadIdx.x >> 6) & 3];
    if (boxed) {
        for (int t = lane; t < B * B * B; t += 64) {
            bx[t] = (uint16_t)label_or_zero(f, x0 + t % B, y0 + (t / B) % B, z0 + t / (B * B));
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
    }
    auto lab_at = [&](int i, int j, int k) {
        return boxed ? (int)bx[(i - x0) + B * ((j - y0) + B * (k - z0))] : label_or_zero(f, i, j, k);
    };
    float best = -1.0f;
    uint64_t lbest = ~0ull;   // not chained: (d^2 bits, scan position) of this lane's best
    float ly[3] = { 0.0f, 0.0f, 0.0f };
    for (int base = 0; base < total; base += 64) {
        const int idx = base + lane;
        int bits = 0;   // bit ax * 4 + e: the face (edge) e of axis ax qualifies
        if (idx < total) {
            const int v[3] = { ci[0] - R + idx % W, ci[1] - R + (idx / W) % W, ci[2] - R + idx / (W * W) };
            if (lab_at(v[0], v[1], v[2]) == a) {
                for (int ax = 0; ax < 3; ++ax) {
                    const int o1 = (ax + 1) % 3, o2 = (ax + 2) % 3;
                    for (int e = 0; e < nsub; ++e) {
                        bool ok;
                        if (c == kNoLab) {
                            int nn[3] = { v[0], v[1], v[2] };
                            nn[ax] += e ? 1 : -1;
                            ok = lab_at(nn[0], nn[1], nn[2]) == b;
                        } else {
                            const int e1 = e & 1, e2 = e >> 1;
                            int hb = 0, hc = 0;
                            for (int m = 0; m < 4; ++m) {
                                int nn[3] = { v[0], v[1], v[2] };
                                nn[o1] += e1 - 1 + (m & 1);
                                nn[o2] += e2 - 1 + (m >> 1);
                                const int ln = lab_at(nn[0], nn[1], nn[2]);
                                hb |= ln == b;
                                hc |= ln == c;
                            }
                            ok = hb && hc;
                        }
                        bits |= ok ? 1 << (ax * 4 + e) : 0;
                        if (ok && !chained) {   // measured from q alone: this lane's first best
                            float y[3];
                            if (c == kNoLab) {
                                y[ax] = (float)(v[ax] + e);
                                y[o1] = fminf(fmaxf(u[o1], (float)v[o1]), (float)(v[o1] + 1));
                                y[o2] = fminf(fmaxf(u[o2], (float)v[o2]), (float)(v[o2] + 1));
                            } else {
                                y[ax] = fminf(fmaxf(u[ax], (float)v[ax]), (float)(v[ax] + 1));
                                y[o1] = (float)(v[o1] + (e & 1));
                                y[o2] = (float)(v[o2] + (e >> 1));
                            }
                            float d2 = 0.0f;
                            for (int k = 0; k < 3; ++k) {
                                y[k] = (y[k] - 0.5f) * vs[k];
                                d2 += (y[k] - q[k]) * (y[k] - q[k]);
                            }
                            const uint64_t key = ((uint64_t)__float_as_uint(d2) << 32) |
                                                 (uint32_t)(idx * 16 + ax * 4 + e);
                            if (key < lbest) {
                                lbest = key;
                                ly[0] = y[0];
                                ly[1] = y[1];
                                ly[2] = y[2];
                            }
                        }
                    }
                }
            }
        }
        if (!chained) {
            continue;
        }
        // the marked voxels of this group in scan order (uniform across the wave)
        uint64_t marked = __ballot(bits != 0);
        while (marked) {
            const int l = __ffsll((long long)marked) - 1;
            marked &= marked - 1;
            const int vb = __shfl(bits, l, 64), vidx = base + l;
            const int v[3] = { ci[0] - R + vidx % W, ci[1] - R + (vidx / W) % W, ci[2] - R + vidx / (W * W) };
            for (int ax = 0; ax < 3; ++ax) {
                const int o1 = (ax + 1) % 3, o2 = (ax + 2) % 3;
                for (int e = 0; e < nsub; ++e) {
                    if (!((vb >> (ax * 4 + e)) & 1)) {
                        continue;
                    }
                    float y[3];
                    if (c == kNoLab) {
                        y[ax] = (float)(v[ax] + e);
                        y[o1] = fminf(fmaxf(u[o1], (float)v[o1]), (float)(v[o1] + 1));
                        y[o2] = fminf(fmaxf(u[o2], (float)v[o2]), (float)(v[o2] + 1));
                    } else {
                        y[ax] = fminf(fmaxf(u[ax], (float)v[ax]), (float)(v[ax] + 1));
                        y[o1] = (float)(v[o1] + (e & 1));
                        y[o2] = (float)(v[o2] + (e >> 1));
                    }
                    float d2 = 0.0f;
                    for (int k = 0; k < 3; ++k) {
                        y[k] = (y[k] - 0.5f) * vs[k];
                        d2 += (y[k] - qc[k]) * (y[k] - qc[k]);
                    }
                    if (best < 0.0f || d2 < best) {
                        best = d2;
                        for (int k = 0; k < 3; ++k) {
                            out[k] = y[k];
                            qc[k] = chained ? y[k] : qc[k];
                        }
                    }
                }
            }
        }
    }
    if (!chained) {   // the smallest (d^2, position): the serial scan's first best
        uint64_t m = lbest;
        for (int d = 32; d > 0; d >>= 1) {
            const uint64_t o = __shfl_xor(m, d, 64);
            m = o < m ? o : m;
        }
        if (m != ~0ull) {
            const int win = __ffsll((long long)__ballot(lbest == m)) - 1;
            out[0] = __shfl(ly[0], win, 64);
            out[1] = __shfl(ly[1], win, 64);
            out[2] = __shfl(ly[2], win, 64);
            best = __uint_as_float((uint32_t)(m >> 32));
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");   // the box is read before it is reused
    __builtin_amdgcn_wave_barrier();
    return best;
}

// v2m_vox_third, src/opencl/v2m_particle_body.cl:533-545
__device__ __forceinline__ int vox_third(const Field& f, int a, int b, const float* p) {
    const int i0 = (int)floorf(over_vx(f, p[0])), j0 = (int)floorf(over_vy(f, p[1])),
              k0 = (int)floorf(over_vz(f, p[2]));
    for (int m = 0; m < 8; ++m) {
        const int l = label_or_zero(f, i0 + (m & 1), j0 + ((m >> 1) & 1), k0 + (m >> 2));
        if (l != a && l != b) {
            return l;
        }
    }
    return kNoLab;
}

// which scans the lane pass leaves to the wave pass: an interior node's wall (bit 0)
// and snap (bit 1) scans, an interface or junction node's face or edge scan (bit 2).
// The snap scans (reach up to 4 voxels, a 9^3 box per label met) are the long ones; the
// face scans (reach about 2) run faster one lane per node (measured: wall and snap
// scans deferred, 5.2 ms an iteration at 2 M nodes; all scans deferred, 6.4 ms).
constexpr int kDeferWall = 1, kDeferSnap = 2, kDeferFace = 4;
#ifndef TN_VOX_DEFER
#define TN_VOX_DEFER 3
#endif
constexpr int kVoxDefer = TN_VOX_DEFER;

// v2m_move_voxel, src/opencl/v2m_particle_body.cl:549-650: the capped step s of a node
// of label a in state ns. Returns |dp| / h (0 and ns untouched when the node's face
// patch or edge vanished within reach). kWave false: one lane per node, and a node
// that needs a face or edge scan (an interior node that hits or nears a wall, every
// interface and junction node) is left untouched, returning kDeferred; kWave true: a
// wave per node (every lane active, the same node), the scans split over the wave.
template <bool kWave, int kDefer = kVoxDefer>
__device__ float voxel_step(const Field& f, float h, float snap, int a, const float* s, NodeState& ns) {
    auto nearest = [&](int bb, int cc, const float* qq, int RR, float* o) {
        return kWave ? vox_nearest_wave(f, a, bb, cc, qq, RR, o) : vox_nearest(f, a, bb, cc, qq, RR, o);
    };
    const float vmin = fminf(fminf(f.vx, f.vy), f.vz);
    const float p[3] = { ns.p.x, ns.p.y, ns.p.z };
    float q[3], x[3];
    int ty = ns.ty, b = ns.b, c = ns.c;
    const float sl = sqrtf(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
    const int R = 1 + (int)ceilf(sl / vmin);
    for (int k = 0; k < 3; ++k) {
        q[k] = p[k] + s[k];
    }
    if (ty == TN_INTERIOR) {
        int l;
        const float t = dda_hit(f, a, p, s, x, &l);
        if (t >= 0.0f) {   // hit a wall: trap on it, glide with the rest of the step
            if (!kWave && (kDefer & kDeferWall)) {
                return kDeferred;
            }
            b = l;
            if (nearest(b, kNoLab, q, R, q) < 0.0f) {
                q[0] = x[0];
                q[1] = x[1];
                q[2] = x[2];
            }
            ty = TN_INTERFACE;
        } else if (snap > 0.0f) {   // close to any wall: snap onto it
            const int Rs = 1 + (int)ceilf(snap * h / vmin);
            const int qx = (int)floorf(over_vx(f, q[0]) + 0.5f), qy = (int)floorf(over_vy(f, q[1]) + 0.5f),
                      qz = (int)floorf(over_vz(f, q[2]) + 0.5f);
            // the label Rs voxels away in direction d (dz, dy, dx = d / 9, d / 3 % 3, d % 3, less 1)
            auto dir_label = [&](int d) {
                return label_or_zero(f, qx + (d % 3 - 1) * Rs, qy + (d / 3 % 3 - 1) * Rs, qz + (d / 9 - 1) * Rs);
            };
            float best = -1.0f, y[3];
            if constexpr (!kWave && (kDefer & kDeferSnap)) {
                // the lane pass hands the node on at the first label other than a and b;
                // its 27 reads are issued at once
                bool other = false;
#pragma unroll
                for (int d = 0; d < 27; ++d) {
                    const int l2 = dir_label(d);
                    other = other | ((l2 != a) & (l2 != b));
                }
                if (other) {
                    return kDeferred;
                }
            } else {
                // the wave reads the 27 labels at once, one a lane
                int lab27 = 0;
                if constexpr (kWave) {
                    const int lane = threadIdx.x & 63;
                    lab27 = dir_label(lane < 27 ? lane : 0);
                }
                // A label met in several directions is scanned once: the scan's result depends
                // on the label only, and a repeat can never beat the best so far (it ties or
                // was rejected), so skipping it keeps the reference's choice.
                constexpr int kSeen = 4;
                int seen[kSeen] = { 0, 0, 0, 0 }, nseen = 0;
                for (int d = 0; d < 27; ++d) {
                    const int l2 = kWave ? __builtin_amdgcn_readlane(lab27, d) : dir_label(d);
                    if (l2 == a || l2 == b) {
                        continue;
                    }
                    bool again = false;
#pragma unroll
                    for (int m = 0; m < kSeen; ++m) {   // (static indices: the list stays in registers)
                        again = again || (m < nseen && seen[m] == l2);
                    }
                    if (again) {
                        continue;
                    }
#pragma unroll
                    for (int m = 0; m < kSeen; ++m) {
                        seen[m] = m == nseen ? l2 : seen[m];
                    }
                    nseen += nseen < kSeen ? 1 : 0;
                    const float d2 = nearest(l2, kNoLab, q, Rs, y);
                    if (d2 >= 0.0f && d2 < snap * snap * h * h && (best < 0.0f || d2 < best)) {
                        best = d2;
                        b = l2;
                        x[0] = y[0];
                        x[1] = y[1];
                        x[2] = y[2];
                    }
                }
            }
            if (best >= 0.0f) {
                q[0] = x[0];
                q[1] = x[1];
                q[2] = x[2];
                ty = TN_INTERFACE;
            }
        }
    } else if (!kWave && (kDefer & kDeferFace)) {
        return kDeferred;
    } else if (ty == TN_INTERFACE) {
        if (nearest(b, kNoLab, q, R, q) < 0.0f) {
            return 0.0f;
        }
    } else {   // junction: the nearest a|b|c voxel edge
        if (nearest(b, c, q, R, q) < 0.0f) {
            return 0.0f;
        }
    }
    // an interface node at a staircase corner where a third label meets: pinned on the
    // nearest a|b|c edge when that is close
    if (ty == TN_INTERFACE) {
        const int cc = vox_third(f, a, b, q);
        if (cc != kNoLab) {
            float r[3];
            const float d2 = nearest(b, cc, q, 1, r);
            if (d2 >= 0.0f && d2 < 0.0625f * h * h) {
                q[0] = r[0];
                q[1] = r[1];
                q[2] = r[2];
                ty = TN_JUNCTION;
                c = cc;
            }
        }
    }
    const float mv = sqrtf((q[0] - p[0]) * (q[0] - p[0]) + (q[1] - p[1]) * (q[1] - p[1]) + (q[2] - p[2]) * (q[2] - p[2]));
    ns.p = p3(q[0], q[1], q[2]);
    ns.ty = ty;
    ns.b = ty == TN_INTERIOR ? kNoLab : b;
    ns.c = ty >= TN_JUNCTION ? c : kNoLab;
    ns.hn = size_at(f, q[0], q[1], q[2]);
    return mv / h;
}

}  // namespace tnd
