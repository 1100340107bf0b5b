// tn_grid.hpp -- the grid build on the device: build_grid_cpu
// (src/v2m_grid.cpp:47-722) for label volumes, tissue-probability inputs and
// gray-scale inputs, as voxel-, brick- and slot-parallel kernels with the host
// driver (tn_grid_build). Included at the end of tn_relax.hip, whose DevMem,
// SeedOps and error handling it uses.
//
// Every per-voxel result is the reference's bit for bit: each sum keeps the
// reference's term order, the Gaussian weights come from the host's expf (the
// reference computes them there too), and the two per-voxel logarithms take the
// host's logf where the device's log could round differently (the thickness
// argument always; a grade only where its bin is in doubt).

namespace tng {

using tnd::Field;
constexpr int kBS = TN_BS;
constexpr int kBL = TN_BL;
constexpr int kSlot = TN_SLOT;
constexpr int kMaxR = 6;

__device__ __forceinline__ int label_at(const Field& f, int i, int j, int k) {
    if (i < 0 || j < 0 || k < 0 || i >= f.nx || j >= f.ny || k >= f.nz) {
        return 0;
    }
    return f.L[tnd::vox(f, i, j, k)];
}

// std::min / std::max on floats as the reference's host code evaluates them
__device__ __forceinline__ float smin(float a, float b) {
    return b < a ? b : a;
}
__device__ __forceinline__ float smax(float a, float b) {
    return a < b ? b : a;
}

// ---- 1. labels near each brick (v2m_brick_labels, src/opencl/v2m_grid_body.cl:28-69) ----
// One workgroup per brick. The labels present in the brick dilated by R (label 0
// outside the grid) are marked in a 65536-bit LDS set; up to kBL of them are the
// set's members in ascending order. A brick with more (an overflow, count kBL + 1)
// keeps the first kBL distinct labels of the reference's z, y, x scan, sorted,
// which one lane finds by that scan.
__global__ __launch_bounds__(256) void k_brick_labels(Field f, int R, int32_t* __restrict__ bl_cnt,
                                                      uint16_t* __restrict__ bl_lab, int32_t* __restrict__ overflow) {
    __shared__ uint32_t bits[2048];
    __shared__ int pref[256];
    const int b = blockIdx.x, tid = threadIdx.x;
    const int bx = b % f.nbx, by = (b / f.nbx) % f.nby, bz = b / (f.nbx * f.nby);
    for (int w = tid; w < 2048; w += 256) {
        bits[w] = 0u;
    }
    __syncthreads();
    const int S = kBS + 2 * R + 1;
    const int x0 = bx * kBS - R, y0 = by * kBS - R, z0 = bz * kBS - R;
    for (int t = tid; t < S * S * S; t += 256) {
        const int l = label_at(f, x0 + t % S, y0 + (t / S) % S, z0 + t / (S * S));
        const uint32_t m = 1u << (l & 31);
        if (!(bits[l >> 5] & m)) {   // most voxels repeat a label already marked
            atomicOr(&bits[l >> 5], m);
        }
    }
    __syncthreads();
    int c = 0;
    for (int w = 8 * tid; w < 8 * tid + 8; ++w) {
        c += __popc(bits[w]);
    }
    pref[tid] = c;
    __syncthreads();
    for (int o = 1; o < 256; o <<= 1) {   // inclusive scan of the counts
        const int v = tid >= o ? pref[tid - o] : 0;
        __syncthreads();
        pref[tid] += v;
        __syncthreads();
    }
    const int total = pref[255];
    uint16_t* out = bl_lab + (size_t)b * kBL;
    if (total <= kBL) {
        int pos = pref[tid] - c;
        for (int w = 8 * tid; w < 8 * tid + 8; ++w) {
            uint32_t m = bits[w];
            while (m) {
                const int q = __ffs(m) - 1;
                m &= m - 1u;
                out[pos++] = (uint16_t)(32 * w + q);
            }
        }
        if (tid >= total && tid < kBL) {
            out[tid] = (uint16_t)0xFFFF;
        }
        if (tid == 0) {
            bl_cnt[b] = total;
        }
        return;
    }
    if (tid == 0) {   // overflow: the reference's scan and insertion
        uint16_t lab[kBL];
        int n = 0;
        for (int z = z0; z < z0 + S; ++z) {
            for (int y = y0; y < y0 + S; ++y) {
                for (int x = x0; x < x0 + S; ++x) {
                    const int l = label_at(f, x, y, z);
                    bool seen = false;
                    for (int i = 0; i < n; ++i) {
                        seen = seen || lab[i] == l;
                    }
                    if (seen || n >= kBL) {
                        continue;
                    }
                    int k = n++;
                    while (k > 0 && lab[k - 1] > l) {
                        lab[k] = lab[k - 1];
                        --k;
                    }
                    lab[k] = (uint16_t)l;
                }
            }
        }
        for (int i = 0; i < kBL; ++i) {
            out[i] = lab[i];
        }
        bl_cnt[b] = kBL + 1;
        atomicAdd(overflow, 1);
    }
}

// ---- 2. slots of the mixed bricks (src/v2m_grid.cpp:126-147) ----
__global__ void k_slot_counts(const int32_t* __restrict__ bl_cnt, int nb, int32_t* __restrict__ cnt) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b < nb) {
        const int n = bl_cnt[b] < kBL ? bl_cnt[b] : kBL;
        cnt[b] = n > 1 ? n : 0;
    }
}

__global__ void k_slot_fill(const int32_t* __restrict__ bl_cnt, const uint16_t* __restrict__ bl_lab, int nb,
                            const int32_t* __restrict__ offs, int32_t* __restrict__ bl_slot,
                            int32_t* __restrict__ slot_brick, uint16_t* __restrict__ slot_label) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) {
        return;
    }
    const int n = bl_cnt[b] < kBL ? bl_cnt[b] : kBL;
    bl_slot[b] = n > 1 ? offs[b] : -1;
    if (n > 1) {
        for (int i = 0; i < n; ++i) {
            slot_brick[offs[b] + i] = b;
            slot_label[offs[b] + i] = bl_lab[(size_t)b * kBL + i];
        }
    }
}

// ---- 3. the smoothed fields of every slot (smooth_all, src/v2m_grid.cpp:153-233) ----
// The source of label l at voxel (x, y, z): its indicator, or (a tissue-probability
// input) its probability, the exterior's indicator outside the grid.
struct Source {
    const float* prob;   // null: the label indicators
    int nprob;
    size_t nvox;
};

__device__ __forceinline__ float source_at(const Field& f, const Source& s, int l, int x, int y, int z) {
    if (!s.prob) {
        return label_at(f, x, y, z) == l ? 1.0f : 0.0f;
    }
    if (x < 0 || y < 0 || z < 0 || x >= f.nx || y >= f.ny || z >= f.nz) {
        return l == 0 ? 1.0f : 0.0f;
    }
    return l < s.nprob ? s.prob[(size_t)l * s.nvox + tnd::vox(f, x, y, z)] : 0.0f;
}

struct Taps {
    float w[2 * kMaxR + 1];   // normalised Gaussian weights, offsets -R..R
    int R;
};

// One workgroup per slot: the (8 + 2R)^3 tile of the slot's brick in LDS, then the
// separable passes of the reference -- x on the centre columns, y on the centre
// rows, z on the brick itself -- each sum over -R..R in that order.
__global__ __launch_bounds__(256) void k_smooth(Field f, Source src, Taps tp, const int32_t* __restrict__ slot_brick,
                                                const uint16_t* __restrict__ slot_label, float* __restrict__ phi) {
    extern __shared__ float tile[];
    const int R = tp.R, T = kBS + 2 * R;
    float* A = tile;
    float* B = tile + T * T * T;
    const int s = blockIdx.x, tid = threadIdx.x;
    const int b = slot_brick[s], l = slot_label[s];
    const int x0 = (b % f.nbx) * kBS - R, y0 = ((b / f.nbx) % f.nby) * kBS - R, z0 = (b / (f.nbx * f.nby)) * kBS - R;
    for (int t = tid; t < T * T * T; t += 256) {
        A[t] = source_at(f, src, l, x0 + t % T, y0 + (t / T) % T, z0 + t / (T * T));
    }
    __syncthreads();
    for (int t = tid; t < T * T * kBS; t += 256) {   // x, centre columns
        const int x = R + t % kBS, y = (t / kBS) % T, z = t / (kBS * T);
        const float* a = A + T * (y + T * z);
        float v = 0.0f;
        for (int dd = -R; dd <= R; ++dd) {
            v += tp.w[dd + R] * a[x + dd];
        }
        B[x + T * (y + T * z)] = v;
    }
    __syncthreads();
    for (int t = tid; t < T * kBS * kBS; t += 256) {   // y, centre rows
        const int x = R + t % kBS, y = R + (t / kBS) % kBS, z = t / (kBS * kBS);
        float v = 0.0f;
        for (int dd = -R; dd <= R; ++dd) {
            v += tp.w[dd + R] * B[x + T * (y + dd + T * z)];
        }
        A[x + T * (y + T * z)] = v;
    }
    __syncthreads();
    for (int t = tid; t < kSlot; t += 256) {   // z, the brick
        const int x = t % kBS + R, y = (t / kBS) % kBS + R, z = t / (kBS * kBS) + R;
        float v = 0.0f;
        for (int dd = -R; dd <= R; ++dd) {
            v += tp.w[dd + R] * A[x + T * (y + T * (z + dd))];
        }
        phi[(size_t)s * kSlot + t] = v;
    }
}

// the raw source as the field (a tissue-probability input with sigma <= 0,
// src/v2m_grid.cpp:434-448)
__global__ void k_source_phi(Field f, Source src, const int32_t* __restrict__ slot_brick,
                             const uint16_t* __restrict__ slot_label, int64_t nslot, float* __restrict__ phi) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= nslot * kSlot) {
        return;
    }
    const int64_t s = g / kSlot;
    const int t = (int)(g % kSlot);
    const int b = slot_brick[s];
    const int bx = b % f.nbx, by = (b / f.nbx) % f.nby, bz = b / (f.nbx * f.nby);
    phi[g] = source_at(f, src, slot_label[s], bx * kBS + t % kBS, by * kBS + (t / kBS) % kBS, bz * kBS + t / (kBS * kBS));
}

// gray-scale memberships (src/v2m_grid.cpp:374-433): phi_l = clamp(0.5 + min(s_{l-1},
// -s_l)) with s_k = (I - t_k) / W_k; outside the grid the exterior's indicator
__global__ void k_gray_phi(Field f, const float* __restrict__ I, const float* __restrict__ TW, int m,
                           const int32_t* __restrict__ slot_brick, const uint16_t* __restrict__ slot_label, int64_t nslot,
                           float* __restrict__ phi) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= nslot * kSlot) {
        return;
    }
    const int64_t s = g / kSlot;
    const int t = (int)(g % kSlot);
    const int b = slot_brick[s], l = slot_label[s];
    const int i = (b % f.nbx) * kBS + t % kBS, j = ((b / f.nbx) % f.nby) * kBS + (t / kBS) % kBS,
              k = (b / (f.nbx * f.nby)) * kBS + t / (kBS * kBS);
    float v;
    if (i >= f.nx || j >= f.ny || k >= f.nz) {
        v = l == 0 ? 1.0f : 0.0f;
    } else {
        const float x = I[tnd::vox(f, i, j, k)];
        float mm = 1e30f;
        if (l >= 1 && l - 1 < m) {
            mm = smin(mm, (x - TW[l - 1]) / TW[m + l - 1]);
        }
        if (l < m) {
            mm = smin(mm, (TW[l] - x) / TW[m + l]);
        }
        v = smin(1.0f, smax(0.0f, 0.5f + mm));
    }
    phi[g] = v;
}

// the intensity steps across threshold t between axis neighbours (the gray-scale
// widths, src/v2m_grid.cpp:382-406), appended as float bits for a sort
__global__ void k_gray_steps(Field f, const float* __restrict__ I, float t, int axis, int32_t* __restrict__ count,
                             int32_t* __restrict__ out) {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nv = (int64_t)f.nx * f.ny * f.nz;
    bool hit = false;
    float d = 0.0f;
    if (v < nv) {
        const int64_t st = axis == 0 ? 1 : (axis == 1 ? (int64_t)f.nx : (int64_t)f.nx * f.ny);
        const int len = axis == 0 ? f.nx : (axis == 1 ? f.ny : f.nz);
        if ((v / st) % len + 1 < len) {
            const float x = I[v] - t, y = I[v + st] - t;
            hit = (x < 0.0f) != (y < 0.0f);
            d = fabsf(y - x);
        }
    }
    const uint64_t m = __ballot(hit);
    if (m) {
        const int lead = __ffsll((long long)m) - 1;
        int base = 0;
        if ((int)(threadIdx.x & 63) == lead) {
            base = atomicAdd(count, __popcll(m));
        }
        base = __shfl(base, lead, 64);
        if (hit) {
            out[base + __popcll(m & tnk::lanemask_lt())] = (int32_t)__float_as_uint(d);
        }
    }
}

// ---- 4. sizing per voxel (v2m_size_voxel / v2m_kappa_max, src/opencl/v2m_grid_body.cl:321-475) ----
struct SizeArgs {
    const float* hlab;
    int nhlab;
    float hbase, hmin, hmax, K, hcurv, iglob;
    const float* ilab;
    int nilab;
    const float* ipair;
    int npair;
    const float* hvox;   // the user's field (> 0 replaces), or null
};

// the largest principal curvature of the level set of phi_a - phi_b through the
// voxel, from central differences at stride 1 (v2m_kappa_max, :321-392)
__device__ __forceinline__ float kappa_max(const Field& f, int a, int b, int i, int j, int k) {
    float F[27];
#pragma unroll
    for (int n = 0; n < 27; ++n) {
        const int x = i + n % 3 - 1, y = j + (n / 3) % 3 - 1, z = k + n / 9 - 1;
        F[n] = tnd::phi_voxel(f, a, x, y, z) - tnd::phi_voxel(f, b, x, y, z);
    }
#define TN_F(dx, dy, dz) F[((dx) + 1) + 3 * ((dy) + 1) + 9 * ((dz) + 1)]
    const float hx = f.vx, hy = f.vy, hz = f.vz;
    float g[3], H[9];
    g[0] = (TN_F(1, 0, 0) - TN_F(-1, 0, 0)) / (2.0f * hx);
    g[1] = (TN_F(0, 1, 0) - TN_F(0, -1, 0)) / (2.0f * hy);
    g[2] = (TN_F(0, 0, 1) - TN_F(0, 0, -1)) / (2.0f * hz);
    H[0] = (TN_F(1, 0, 0) - 2.0f * TN_F(0, 0, 0) + TN_F(-1, 0, 0)) / (hx * hx);
    H[4] = (TN_F(0, 1, 0) - 2.0f * TN_F(0, 0, 0) + TN_F(0, -1, 0)) / (hy * hy);
    H[8] = (TN_F(0, 0, 1) - 2.0f * TN_F(0, 0, 0) + TN_F(0, 0, -1)) / (hz * hz);
    H[1] = H[3] = (TN_F(1, 1, 0) - TN_F(1, -1, 0) - TN_F(-1, 1, 0) + TN_F(-1, -1, 0)) / (4.0f * hx * hy);
    H[2] = H[6] = (TN_F(1, 0, 1) - TN_F(1, 0, -1) - TN_F(-1, 0, 1) + TN_F(-1, 0, -1)) / (4.0f * hx * hz);
    H[5] = H[7] = (TN_F(0, 1, 1) - TN_F(0, 1, -1) - TN_F(0, -1, 1) + TN_F(0, -1, -1)) / (4.0f * hy * hz);
#undef TN_F
    const float gn = sqrtf(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
    if (gn < 1e-6f) {
        return 0.0f;
    }
    const float n[3] = { g[0] / gn, g[1] / gn, g[2] / gn };
    float P[9], HP[9];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            P[3 * r + c] = (r == c ? 1.0f : 0.0f) - n[r] * n[c];
        }
    }
#pragma unroll
    for (int r = 0; r < 3; ++r) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            HP[3 * r + c] = H[3 * r] * P[c] + H[3 * r + 1] * P[3 + c] + H[3 * r + 2] * P[6 + c];
        }
    }
    float tr = 0.0f, fro = 0.0f, diag[3];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const float q = P[3 * r] * HP[c] + P[3 * r + 1] * HP[3 + c] + P[3 * r + 2] * HP[6 + c];
            fro += q * q;
            if (r == c) {
                diag[r] = q;
            }
        }
    }
    tr = (diag[0] + diag[1] + diag[2]) / gn;
    fro /= gn * gn;
    const float disc = 2.0f * fro - tr * tr;
    return 0.5f * (fabsf(tr) + sqrtf(disc > 0.0f ? disc : 0.0f));
}

// v2m_isize_of (:394-411): the a|b interface size
__device__ __forceinline__ float isize_of(const SizeArgs& z, int a, int b) {
    const float lo = (float)(a < b ? a : b), hi = (float)(a < b ? b : a);
    for (int p = 0; p < z.npair; ++p) {
        if (z.ipair[3 * p] == lo && z.ipair[3 * p + 1] == hi) {
            return z.ipair[3 * p + 2];
        }
    }
    const float ha = a < z.nilab ? z.ilab[a] : 0.0f, hb = b < z.nilab ? z.ilab[b] : 0.0f;
    if (ha > 0.0f && hb > 0.0f) {
        return fminf(ha, hb);
    }
    return ha > 0.0f ? ha : (hb > 0.0f ? hb : z.iglob);
}

// v2m_size_voxel (:413-475), then the user's field where it is > 0 (src/v2m_grid.cpp:282-290)
__global__ void k_size(Field f, SizeArgs z, float* __restrict__ h) {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nv = (int64_t)f.nx * f.ny * f.nz;
    if (v >= nv) {
        return;
    }
    const int i = (int)(v % f.nx), j = (int)((v / f.nx) % f.ny), k = (int)(v / ((int64_t)f.nx * f.ny));
    const int a = f.L[v];
    const bool isz = z.iglob > 0.0f || z.nilab > 0 || z.npair > 0;
    float hv = (a < z.nhlab && z.hlab[a] > 0.0f) ? z.hlab[a] : z.hbase;
    if (a == 0) {
        hv = z.hmax;
    } else {
        const int bk = tnd::brick_at(f, i, j, k);
        if (f.bl_slot[bk] >= 0) {
            const float pa = tnd::phi_voxel(f, a, i, j, k);
            int best = -1;
            float pb = -1.0f;
            const int cnt = f.bl_cnt[bk] < kBL ? f.bl_cnt[bk] : kBL;
            for (int s = 0; s < cnt; ++s) {
                const int l = f.bl_lab[(size_t)bk * kBL + s];
                if (l == a) {
                    continue;
                }
                const float pv = tnd::phi_voxel(f, l, i, j, k);
                if (pv > pb) {
                    pb = pv;
                    best = l;
                }
                if (isz && pa - pv < 0.45f) {
                    const float hi = isize_of(z, a, l);
                    if (hi > 0.0f) {
                        hv = fminf(hv, hi);
                    }
                }
            }
            if (best >= 0 && pa - pb < 0.45f) {
                if (best < z.nhlab && z.hlab[best] > 0.0f) {
                    hv = fminf(hv, z.hlab[best]);
                }
                const float kap = kappa_max(f, a, best, i, j, k);
                if (kap > 0.0f) {
                    hv = fminf(hv, fmaxf(1.0f / (z.K * kap), z.hcurv));
                }
            }
        }
        hv = hv < z.hmin ? z.hmin : (hv > z.hmax ? z.hmax : hv);
    }
    if (z.hvox && z.hvox[v] > 0.0f) {
        hv = z.hvox[v];
    }
    h[v] = hv;
}

// ---- thickness (v2m_thick_voxel, :241-276) ----
// Pass 1 (k_own_phi, k_thick_window below): m = the largest phi of the voxel's own
// label over its label's voxels within +-2; where the thickness is finite, (voxel,
// m, (1 - m)(1 + m)) is listed for the logarithm. Elsewhere tvox = 1e30.
// Pass 2: erfinv (M. Giles, "Approximating the erfinv function", GPU Computing Gems
// Jade, 2011; single-precision coefficients) from the host's logarithm, and
// t = 2 sqrt(2) sigma erfinv(m)
__global__ void k_thick_finish(int32_t n, const int32_t* __restrict__ idx, const float* __restrict__ mval,
                               const float* __restrict__ lg, float sigma, float* __restrict__ tvox) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) {
        return;
    }
    const float x = mval[q];
    float w = -lg[q], p;
    if (w < 5.0f) {
        w -= 2.5f;
        p = 2.81022636e-08f;
        p = 3.43273939e-07f + p * w;
        p = -3.5233877e-06f + p * w;
        p = -4.39150654e-06f + p * w;
        p = 0.00021858087f + p * w;
        p = -0.00125372503f + p * w;
        p = -0.00417768164f + p * w;
        p = 0.246640727f + p * w;
        p = 1.50140941f + p * w;
    } else {
        w = sqrtf(w) - 3.0f;
        p = -0.000200214257f;
        p = 0.000100950558f + p * w;
        p = 0.00134934322f + p * w;
        p = -0.00367342844f + p * w;
        p = 0.00573950773f + p * w;
        p = -0.0076224613f + p * w;
        p = 0.00943887047f + p * w;
        p = 1.00167406f + p * w;
        p = 2.83297682f + p * w;
    }
    tvox[idx[q]] = 2.8284271f * sigma * (p * x);
}

// thin layers: h <= t vmin / thick, no smaller than floor_mm (src/v2m_grid.cpp:318-337)
__global__ void k_thin_size(int64_t nv, const float* __restrict__ tvox, float vmin, float thick, float floor_mm,
                            float* __restrict__ h) {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v < nv) {
        const float t = tvox[v];
        if (t < 1e29f) {
            h[v] = smin(h[v], smax(floor_mm, t * vmin / thick));
        }
    }
}

// ---- the thickness-adaptive blend (src/v2m_grid.cpp:472-541) ----
// one separable pass along an axis with clamped borders: the minimum over +-r, or
// the weighted sum over -r..r in that order
__global__ void k_axis_pass(Field f, int axis, int r, int mn, Taps tp, const float* __restrict__ src,
                            float* __restrict__ dst) {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nv = (int64_t)f.nx * f.ny * f.nz;
    if (v >= nv) {
        return;
    }
    const int64_t st = axis == 0 ? 1 : (axis == 1 ? (int64_t)f.nx : (int64_t)f.nx * f.ny);
    const int len = axis == 0 ? f.nx : (axis == 1 ? f.ny : f.nz);
    const int c = (int)((v / st) % len);
    float acc = mn ? 1e30f : 0.0f;
    for (int dd = -r; dd <= r; ++dd) {
        const int cc = min(len - 1, max(0, c + dd));
        const float x = src[v + (int64_t)(cc - c) * st];
        acc = mn ? smin(acc, x) : acc + tp.w[dd + r] * x;
    }
    dst[v] = acc;
}

__global__ void k_smoothstep(int64_t nv, float lo, float den, float* __restrict__ w) {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v < nv) {
        float x = (w[v] - lo) / den;
        x = smin(1.0f, smax(0.0f, x));
        w[v] = x * x * (3.0f - 2.0f * x);
    }
}

// phi = w phi_smooth + (1 - w) phi_sharp inside the grid; slot entries past the
// grid's edge keep the sharp value
__global__ void k_blend(Field f, const int32_t* __restrict__ slot_brick, int64_t nslot, const float* __restrict__ w,
                        const float* __restrict__ phi_s, float* __restrict__ phi) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= nslot * kSlot) {
        return;
    }
    const int b = slot_brick[g / kSlot];
    const int t = (int)(g % kSlot);
    const int i = (b % f.nbx) * kBS + t % kBS, j = ((b / f.nbx) % f.nby) * kBS + (t / kBS) % kBS,
              k = (b / (f.nbx * f.nby)) * kBS + t / (kBS * kBS);
    if (i >= f.nx || j >= f.ny || k >= f.nz) {
        return;
    }
    const float wv = w[tnd::vox(f, i, j, k)];
    phi[g] = wv * phi_s[g] + (1.0f - wv) * phi[g];
}

// ---- the label-preserving margin (v2m_preserve_voxel, :282-317) ----
__global__ void k_preserve(Field f, float mg, float* __restrict__ phi) {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nv = (int64_t)f.nx * f.ny * f.nz;
    if (v >= nv) {
        return;
    }
    const int i = (int)(v % f.nx), j = (int)((v / f.nx) % f.ny), k = (int)(v / ((int64_t)f.nx * f.ny));
    const int b = tnd::brick_at(f, i, j, k);
    const int s0 = f.bl_slot[b];
    if (s0 < 0) {
        return;
    }
    const int a = f.L[v];
    const size_t t = (size_t)((i & 7) + kBS * ((j & 7) + kBS * (k & 7)));
    const int n = f.bl_cnt[b] < kBL ? f.bl_cnt[b] : kBL;
    int sa = -1;
    float mo = 0.0f;
    for (int s = 0; s < n; ++s) {
        const float q = phi[(size_t)(s0 + s) * kSlot + t];
        if (f.bl_lab[(size_t)b * kBL + s] == a) {
            sa = s0 + s;
        } else if (q > mo) {
            mo = q;
        }
    }
    if (sa >= 0 && phi[(size_t)sa * kSlot + t] < mo + mg) {
        phi[(size_t)sa * kSlot + t] = mo + mg;
    }
}

// ---- 5. gradient limiting: Jacobi sweeps (v2m_limit_voxel, :477-496) ----
// h <= h(n) + g |x - x_n| over the 26 neighbours, every voxel from the previous
// sweep's values, until no value drops.
// One sweep over listed bricks, one workgroup of 512 per brick: the brick and a
// one-voxel halo in LDS (+inf outside the grid, which no minimum takes), so each
// voxel reads its 26 neighbours from LDS; the neighbour distances g |dx| are formed
// once per workgroup. A minimum is exact in any order, so the result is the
// reference's sweep. A brick is listed while a brick within one of it changed in the
// sweep before (the others cannot change), which keeps the reference's count.
__global__ __launch_bounds__(512) void k_limit_bricks(Field f, float gl, const int32_t* __restrict__ list,
                                                      const float* __restrict__ h, float* __restrict__ hn,
                                                      int32_t* __restrict__ bchanged) {
    __shared__ float tile[10 * 10 * 10];
    __shared__ float dist[27];
    const int b = list[blockIdx.x], tid = threadIdx.x;
    const int bx = b % f.nbx, by = (b / f.nbx) % f.nby, bz = b / (f.nbx * f.nby);
    const int x0 = bx * kBS - 1, y0 = by * kBS - 1, z0 = bz * kBS - 1;
    for (int t = tid; t < 1000; t += 512) {
        const int x = x0 + t % 10, y = y0 + (t / 10) % 10, z = z0 + t / 100;
        tile[t] = (x < 0 || y < 0 || z < 0 || x >= f.nx || y >= f.ny || z >= f.nz) ? INFINITY
                                                                                    : h[tnd::vox(f, x, y, z)];
    }
    if (tid < 27) {
        const float ex = (float)(tid % 3 - 1) * f.vx, ey = (float)((tid / 3) % 3 - 1) * f.vy,
                    ez = (float)(tid / 9 - 1) * f.vz;
        dist[tid] = gl * sqrtf(ex * ex + ey * ey + ez * ez);
    }
    __syncthreads();
    const int lx = tid % kBS, ly = (tid / kBS) % kBS, lz = tid / (kBS * kBS);
    const int i = bx * kBS + lx, j = by * kBS + ly, k = bz * kBS + lz;
    if (i >= f.nx || j >= f.ny || k >= f.nz) {
        return;
    }
    const float* c = tile + (lx + 1) + 10 * ((ly + 1) + 10 * (lz + 1));
    const float h0 = c[0];
    float m = h0;
#pragma unroll
    for (int n = 0; n < 27; ++n) {
        if (n != 13) {
            m = fminf(m, c[(n % 3 - 1) + 10 * ((n / 3) % 3 - 1) + 100 * (n / 9 - 1)] + dist[n]);
        }
    }
    const size_t v = tnd::vox(f, i, j, k);
    hn[v] = m;
    if (m < h0) {
        bchanged[b] = 1;
    }
}

// h = hn on the listed bricks
__global__ __launch_bounds__(512) void k_commit_bricks(Field f, const int32_t* __restrict__ list,
                                                       const float* __restrict__ hn, float* __restrict__ h) {
    const int b = list[blockIdx.x], tid = threadIdx.x;
    const int i = (b % f.nbx) * kBS + tid % kBS, j = ((b / f.nbx) % f.nby) * kBS + (tid / kBS) % kBS,
              k = (b / (f.nbx * f.nby)) * kBS + tid / (kBS * kBS);
    if (i < f.nx && j < f.ny && k < f.nz) {
        const size_t v = tnd::vox(f, i, j, k);
        h[v] = hn[v];
    }
}

// the bricks of the next sweep: those with a changed brick within one; the flags
// are read from bchanged and cleared in bnext's twin by the caller
__global__ void k_next_bricks(Field f, const int32_t* __restrict__ bchanged, int32_t* __restrict__ count,
                              int32_t* __restrict__ list) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    const int nb = f.nbx * f.nby * f.nbz;
    bool on = false;
    if (b < nb) {
        const int bx = b % f.nbx, by = (b / f.nbx) % f.nby, bz = b / (f.nbx * f.nby);
        for (int dz = -1; dz <= 1 && !on; ++dz) {
            for (int dy = -1; dy <= 1 && !on; ++dy) {
                for (int dx = -1; dx <= 1 && !on; ++dx) {
                    const int x = bx + dx, y = by + dy, z = bz + dz;
                    on = x >= 0 && y >= 0 && z >= 0 && x < f.nbx && y < f.nby && z < f.nbz &&
                         bchanged[x + f.nbx * (y + f.nby * z)] != 0;
                }
            }
        }
    }
    const uint64_t bal = __ballot(on);
    if (bal) {
        const int lead = __ffsll((long long)bal) - 1;
        int base = 0;
        if ((int)(threadIdx.x & 63) == lead) {
            base = atomicAdd(count, __popcll(bal));
        }
        base = __shfl(base, lead, 64);
        if (on) {
            list[base + __popcll(bal & tnk::lanemask_lt())] = b;
        }
    }
}

// Only voxels of label a enter voxel v's maximum, each with phi_a at itself: so m
// is the largest "own phi" (phi of a voxel's own label at the voxel, 1 for the
// exterior outside the grid) among the voxels of label a within +-2. k_own_phi forms
// the own phi per voxel; k_thick_window takes the window maximum from an LDS tile of the
// brick and a two-voxel halo.
__global__ void k_own_phi(Field f, float* __restrict__ own) {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nv = (int64_t)f.nx * f.ny * f.nz;
    if (v < nv) {
        const int i = (int)(v % f.nx), j = (int)((v / f.nx) % f.ny), k = (int)(v / ((int64_t)f.nx * f.ny));
        own[v] = tnd::phi_voxel(f, f.L[v], i, j, k);
    }
}

__global__ __launch_bounds__(512) void k_thick_window(Field f, const float* __restrict__ own, float* __restrict__ tvox,
                                                      int32_t* __restrict__ count, int32_t* __restrict__ idx,
                                                      float* __restrict__ mval, float* __restrict__ larg) {
    constexpr int W = kBS + 4;
    __shared__ float tp[W * W * W];
    __shared__ uint16_t tl[W * W * W];
    const int b = blockIdx.x, tid = threadIdx.x;
    const int bx = b % f.nbx, by = (b / f.nbx) % f.nby, bz = b / (f.nbx * f.nby);
    const int lx = tid % kBS, ly = (tid / kBS) % kBS, lz = tid / (kBS * kBS);
    const int i = bx * kBS + lx, j = by * kBS + ly, k = bz * kBS + lz;
    const bool in = i < f.nx && j < f.ny && k < f.nz;
    if (f.bl_slot[b] < 0) {   // a uniform brick: no finite thickness
        if (in) {
            tvox[tnd::vox(f, i, j, k)] = 1e30f;
        }
        return;
    }
    const int x0 = bx * kBS - 2, y0 = by * kBS - 2, z0 = bz * kBS - 2;
    for (int t = tid; t < W * W * W; t += 512) {
        const int x = x0 + t % W, y = y0 + (t / W) % W, z = z0 + t / (W * W);
        const bool ok = x >= 0 && y >= 0 && z >= 0 && x < f.nx && y < f.ny && z < f.nz;
        const size_t v = ok ? tnd::vox(f, x, y, z) : 0;
        tl[t] = ok ? f.L[v] : (uint16_t)0;
        tp[t] = ok ? own[v] : 1.0f;
    }
    __syncthreads();
    bool want = false;
    float m = 0.0f;
    if (in) {
        const int c = (lx + 2) + W * ((ly + 2) + W * (lz + 2));
        const int a = tl[c];
        for (int dz = -2; dz <= 2; ++dz) {
            for (int dy = -2; dy <= 2; ++dy) {
                for (int dx = -2; dx <= 2; ++dx) {
                    const int q = c + dx + W * (dy + W * dz);
                    if (tl[q] == a) {
                        m = tp[q] > m ? tp[q] : m;
                    }
                }
            }
        }
        want = !(m >= 0.9999f);
        tvox[tnd::vox(f, i, j, k)] = 1e30f;
    }
    const uint64_t bal = __ballot(want);
    if (bal) {
        const int lead = __ffsll((long long)bal) - 1;
        int base = 0;
        if ((int)(threadIdx.x & 63) == lead) {
            base = atomicAdd(count, __popcll(bal));
        }
        base = __shfl(base, lead, 64);
        if (want) {
            const int at = base + __popcll(bal & tnk::lanemask_lt());
            idx[at] = (int32_t)tnd::vox(f, i, j, k);
            mval[at] = m;
            larg[at] = (1.0f - m) * (1.0f + m);
        }
    }
}

// ---- 6. grades (v2m_grade, :498-506) ----
// grade = (uint8)(t 255 + 0.5), t = clamp(log(h / hmin) / log(hmax / hmin)). The
// numerator's logarithm is the device's double-precision one rounded to float; the
// host's logf is within one ulp of it, so the grade is settled when all three
// neighbouring floats give the same bin. The few voxels where they do not are
// listed for the host's logf.
__global__ void k_grade(int64_t nv, const float* __restrict__ h, float hmin, float lden, int32_t* __restrict__ count,
                        int32_t* __restrict__ doubt, uint8_t* __restrict__ grade) {
    const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool open = false;
    if (v < nv) {
        const float q = h[v] / hmin;
        const float c = (float)log((double)q);
        const float cand[3] = { nextafterf(c, -INFINITY), c, nextafterf(c, INFINITY) };
        int gb[3];
#pragma unroll
        for (int u = 0; u < 3; ++u) {
            float t = cand[u] / lden;
            t = t < 0.0f ? 0.0f : (t > 1.0f ? 1.0f : t);
            gb[u] = (int)(t * 255.0f + 0.5f);
        }
        open = gb[0] != gb[1] || gb[1] != gb[2];
        grade[v] = (uint8_t)gb[1];
    }
    const uint64_t bal = __ballot(open);
    if (bal) {
        const int lead = __ffsll((long long)bal) - 1;
        int base = 0;
        if ((int)(threadIdx.x & 63) == lead) {
            base = atomicAdd(count, __popcll(bal));
        }
        base = __shfl(base, lead, 64);
        if (open) {
            doubt[base + __popcll(bal & tnk::lanemask_lt())] = (int32_t)v;
        }
    }
}

// the sizes of the listed voxels, and their grades put back
__global__ void k_gather_f32(int32_t n, const int32_t* __restrict__ idx, const float* __restrict__ a,
                             float* __restrict__ out) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) {
        out[q] = a[idx[q]];
    }
}

__global__ void k_scatter_u8(int32_t n, const int32_t* __restrict__ idx, const uint8_t* __restrict__ val,
                             uint8_t* __restrict__ a) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) {
        a[idx[q]] = val[q];
    }
}

__global__ void k_fill_u8(uint8_t* __restrict__ a, int64_t n, uint8_t x) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        a[i] = x;
    }
}

__global__ void k_pos_minmax(const float* __restrict__ a, int64_t n, int32_t* __restrict__ mm) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && a[i] > 0.0f) {   // positive floats order as their bits
        atomicMin(&mm[0], (int32_t)__float_as_uint(a[i]));
        atomicMax(&mm[1], (int32_t)__float_as_uint(a[i]));
    }
}

}  // namespace tng

// ---- the host driver ------------------------------------------------------------------

struct tn_grid {
    DevMem mem;
    tn_grid_info info{};
    tnd::Field f{};                 // the grid's device arrays (L, bricks, slots, phi, h)
    uint8_t* grade = nullptr;       // [nx*ny*nz]
    std::vector<float> gTW;         // gray-scale: the thresholds, then their widths
    const float* gI = nullptr;      // gray-scale: the intensity on the device
    const float* dTW = nullptr;     // gray-scale: gTW on the device
    float hbig = 0.0f;              // the largest h
    // the last tn_seed_count_from_grid's nodes, kept on the device for the
    // tn_seed_from_grid that follows it with the same parameters
    struct SeedCache {
        DevMem mem;
        tn_relax_params prm{};
        Seeds s;
    };
    mutable std::unique_ptr<SeedCache> seeds;
};

namespace {

// v2m_gauss's taps (src/opencl/v2m_grid_body.cl:81-84), normalised as smooth_all
// does (src/v2m_grid.cpp:171-179): the host's expf, like the reference's
tng::Taps gauss_taps(float sigma, int R) {
    tng::Taps t;
    std::memset(&t, 0, sizeof t);
    t.R = R;
    float ws = 0.0f;
    for (int d = -R; d <= R; ++d) {
        t.w[d + R] = std::exp(-0.5f * (float)(d * d) / (sigma * sigma));
        ws += t.w[d + R];
    }
    for (int d = 0; d <= 2 * R; ++d) {
        t.w[d] /= ws;
    }
    return t;
}

int smooth_radius(float sigma) {
    return std::min(6, std::max(1, (int)std::ceil(3.0f * sigma)));
}

void check_volume(const tn_volume_view* v, const tn_grid_params* p) {
    arg_check(v && p, "null argument");
    arg_check(v->nx > 0 && v->ny > 0 && v->nz > 0, "volume: empty");
    arg_check((int64_t)v->nx * v->ny * v->nz < ((int64_t)1 << 31), "volume: more than 2^31 voxels");
    arg_check(v->vs[0] > 0 && v->vs[1] > 0 && v->vs[2] > 0, "volume: voxel sizes must be positive");
    arg_check(v->L != nullptr, "volume: null labels");
    arg_check(v->nprob >= 0 && (v->nprob == 0 || v->prob), "volume: nprob > 0 needs prob");
    arg_check(v->nthr >= 0 && (v->nthr == 0 || (v->gray && v->thresholds)), "volume: nthr > 0 needs gray and thresholds");
    for (int k = 1; k < v->nthr; ++k) {
        arg_check(v->thresholds[k - 1] <= v->thresholds[k], "volume: thresholds must ascend");
    }
    arg_check(p->sigma_curv > 0 && p->K > 0 && p->g >= 0, "grid params: need sigma_curv > 0, K > 0, g >= 0");
    arg_check(p->nhlab >= 0 && (p->nhlab == 0 || p->hlab), "grid params: nhlab > 0 needs hlab");
    arg_check(p->nilab >= 0 && (p->nilab == 0 || p->ilab), "grid params: nilab > 0 needs ilab");
    arg_check(p->npair >= 0 && (p->npair == 0 || p->ipair), "grid params: npair > 0 needs ipair");
}

// the host's logf over a list (the thickness arguments), on a few threads
void host_logf(std::vector<float>& a) {
    const size_t n = a.size();
    const unsigned nt = (unsigned)std::max<size_t>(1, std::min<size_t>(16, n / 65536));
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; ++t) {
        th.emplace_back([&a, n, nt, t] {
            for (size_t i = t * n / nt; i < (t + 1) * n / nt; ++i) {
                a[i] = std::log(a[i]);
            }
        });
    }
    for (auto& x : th) {
        x.join();
    }
}

}  // namespace

extern "C" {

void tn_grid_params_default(tn_grid_params* p) {
    if (!p) {
        return;
    }
    std::memset(p, 0, sizeof *p);
    p->sigma = 1.0f;
    p->sigma_thin = 0.35f;
    p->thin_lo = 2.0f;
    p->thin_hi = 4.0f;
    p->thick = 0.0f;
    p->thin_floor = 0.5f;
    p->preserve = 0.0f;
    p->sigma_curv = 2.0f;
    p->K = 3.0f;
    p->g = 0.3f;
}

int tn_grid_build(const tn_volume_view* vol, const tn_grid_params* prm, int32_t device, tn_grid** out) {
    return guarded([&] {
        arg_check(out != nullptr, "null output handle");
        *out = nullptr;
        check_volume(vol, prm);
        const int dev = pick_device(device);
        TN_HIP(hipSetDevice(dev));
        std::unique_ptr<tn_grid> G(new tn_grid);
        G->mem.device = dev;
        DevMem& mem = G->mem;
        DevMem tmp;   // scratch freed at the end of the build
        tmp.device = dev;
        SeedOps ops{ tmp };
        // TN_GRID_VERBOSE: device time per stage (events on the null stream)
        const bool verbose = std::getenv("TN_GRID_VERBOSE") != nullptr;
        std::vector<std::pair<const char*, hipEvent_t>> marks;
        auto mark = [&](const char* name) {
            if (verbose) {
                hipEvent_t e;
                TN_HIP(hipEventCreate(&e));
                TN_HIP(hipEventRecord(e, nullptr));
                marks.emplace_back(name, e);
            }
        };
        const int nx = vol->nx, ny = vol->ny, nz = vol->nz;
        const int64_t nv = (int64_t)nx * ny * nz;
        tn_grid_info& I = G->info;
        I.nx = nx;
        I.ny = ny;
        I.nz = nz;
        I.nbx = (nx + TN_BS - 1) / TN_BS;
        I.nby = (ny + TN_BS - 1) / TN_BS;
        I.nbz = (nz + TN_BS - 1) / TN_BS;
        const int nb = I.nbx * I.nby * I.nbz;
        // sizes and radii (src/v2m_grid.cpp:56-95)
        const int Ri = smooth_radius(prm->sigma), Rc = smooth_radius(prm->sigma_curv);
        I.R = std::max(Ri, Rc);
        const float vmin = std::min(vol->vs[0], std::min(vol->vs[1], vol->vs[2]));
        float hbase = prm->hbase > 0 ? prm->hbase : 3.0f * vmin;
        float hmin = prm->hmin > 0 ? prm->hmin : hbase / 3.0f;
        float hmax = prm->hmax > 0 ? prm->hmax : hbase;
        for (int l = 0; l < prm->nhlab; ++l) {
            if (prm->hlab[l] > 0) {
                hmax = std::max(hmax, prm->hlab[l]);
                hmin = std::min(hmin, prm->hlab[l]);
            }
        }
        const float hcurv = hmin;
        auto span = [&](float v) {
            if (v > 0) {
                hmin = std::min(hmin, v);
                hmax = std::max(hmax, v);
            }
        };
        span(prm->iglob);
        for (int l = 0; l < prm->nilab; ++l) {
            span(prm->ilab[l]);
        }
        for (int p = 0; p < prm->npair; ++p) {
            span(prm->ipair[3 * p + 2]);
        }
        const unsigned nbv = blocks_for(nv, 256);
        const float* d_hvox = nullptr;
        if (vol->hvox) {   // the user's field: the grades span its sizes too
            d_hvox = tmp.upload(vol->hvox, (size_t)nv);
            int32_t* mm = tmp.alloc<int32_t>(2);
            const int32_t init[2] = { 0x7F800000, 0 };
            TN_HIP(hipMemcpy(mm, init, sizeof init, hipMemcpyHostToDevice));
            tng::k_pos_minmax<<<nbv, 256>>>(d_hvox, nv, mm);
            int32_t got[2];
            TN_HIP(hipMemcpy(got, mm, sizeof got, hipMemcpyDeviceToHost));
            if (got[1] > 0) {
                float lo, hi;
                std::memcpy(&lo, &got[0], 4);
                std::memcpy(&hi, &got[1], 4);
                hmax = std::max(hmax, hi);
                hmin = std::min(hmin, lo);
            }
        }
        tnd::Field& f = G->f;
        std::memset(&f, 0, sizeof f);
        f.nx = nx;
        f.ny = ny;
        f.nz = nz;
        f.nbx = I.nbx;
        f.nby = I.nby;
        f.nbz = I.nbz;
        f.vx = vol->vs[0];
        f.vy = vol->vs[1];
        f.vz = vol->vs[2];
        f.L = mem.upload(vol->L, (size_t)nv);
        mark("upload");
        // 1. bricks
        int32_t* bl_cnt = mem.alloc<int32_t>(nb);
        uint16_t* bl_lab = mem.alloc<uint16_t>((size_t)nb * TN_BL);
        int32_t* ovf = ops.counts(0);
        tng::k_brick_labels<<<nb, 256>>>(f, I.R, bl_cnt, bl_lab, ovf);
        TN_HIP(hipGetLastError());
        mark("bricks");
        // 2. slots
        int32_t* offs = ops.counts(nb);
        tng::k_slot_counts<<<blocks_for(nb, 256), 256>>>(bl_cnt, nb, offs);
        const int32_t nslot = ops.scan(offs, nb);
        int32_t* bl_slot = mem.alloc<int32_t>(nb);
        int32_t* slot_brick = tmp.alloc<int32_t>(std::max(1, nslot));
        uint16_t* slot_label = tmp.alloc<uint16_t>(std::max(1, nslot));
        tng::k_slot_fill<<<blocks_for(nb, 256), 256>>>(bl_cnt, bl_lab, nb, offs, bl_slot, slot_brick, slot_label);
        TN_HIP(hipMemcpy(&I.overflow_bricks, ovf, sizeof(int32_t), hipMemcpyDeviceToHost));
        I.nslot = nslot;
        f.bl_cnt = bl_cnt;
        f.bl_lab = bl_lab;
        f.bl_slot = bl_slot;
        const size_t nphi = (size_t)nslot * TN_SLOT;
        float* phi = mem.alloc<float>(std::max<size_t>(1, nphi));
        f.phi = phi;
        const tng::Source labels_src{ nullptr, 0, (size_t)nv };
        const bool tpm = vol->prob && vol->nprob > 0;
        const bool gray = !tpm && vol->nthr > 0;
        I.prob_fields = tpm ? 1 : 0;   // Grid::prob_fields (src/v2m_grid.cpp:343-345)
        const tng::Source src{ tpm ? tmp.upload(vol->prob, (size_t)nv * vol->nprob) : nullptr, tpm ? vol->nprob : 0,
                               (size_t)nv };
        auto smooth = [&](const tng::Source& s, float sigma, int R, float* dst) {
            if (nslot == 0) {
                return;
            }
            const int T = TN_BS + 2 * R;
            tng::k_smooth<<<nslot, 256, 2 * T * T * T * sizeof(float)>>>(f, s, gauss_taps(sigma, R), slot_brick,
                                                                        slot_label, dst);
            TN_HIP(hipGetLastError());
        };
        mark("slots");
        // 3. the curvature fields, from the labels
        smooth(labels_src, prm->sigma_curv, Rc, phi);
        mark("smooth_curv");
        // 4. sizing
        float* h = mem.alloc<float>((size_t)nv);
        tng::SizeArgs za;
        std::memset(&za, 0, sizeof za);
        za.hlab = prm->nhlab ? tmp.upload(prm->hlab, (size_t)prm->nhlab) : nullptr;
        za.nhlab = prm->nhlab;
        za.hbase = hbase;
        za.hmin = hmin;
        za.hmax = hmax;
        za.K = prm->K;
        za.hcurv = hcurv;
        za.iglob = prm->iglob;
        za.ilab = prm->nilab ? tmp.upload(prm->ilab, (size_t)prm->nilab) : nullptr;
        za.nilab = prm->nilab;
        za.ipair = prm->npair ? tmp.upload(prm->ipair, (size_t)prm->npair * 3) : nullptr;
        za.npair = prm->npair;
        za.hvox = d_hvox;
        tng::k_size<<<nbv, 256>>>(f, za, h);
        TN_HIP(hipGetLastError());
        mark("size");
        // thickness from the curvature fields (src/v2m_grid.cpp:292-305)
        float* tvox = nullptr;
        if (prm->thick > 0.0f || prm->sigma_thin > 0.0f) {
            tvox = tmp.alloc<float>((size_t)nv);
            int32_t* cnt = ops.counts(0);
            int32_t* idx = tmp.alloc<int32_t>((size_t)nv);
            float* mval = tmp.alloc<float>((size_t)nv);
            float* larg = tmp.alloc<float>((size_t)nv);
            float* own = tmp.alloc<float>((size_t)nv);
            tng::k_own_phi<<<nbv, 256>>>(f, own);
            tng::k_thick_window<<<nb, 512>>>(f, own, tvox, cnt, idx, mval, larg);
            TN_HIP(hipGetLastError());
            int32_t n = 0;
            TN_HIP(hipMemcpy(&n, cnt, sizeof n, hipMemcpyDeviceToHost));
            if (n > 0) {
                std::vector<float> lg((size_t)n);
                TN_HIP(hipMemcpy(lg.data(), larg, sizeof(float) * n, hipMemcpyDeviceToHost));
                host_logf(lg);
                TN_HIP(hipMemcpy(larg, lg.data(), sizeof(float) * n, hipMemcpyHostToDevice));
                tng::k_thick_finish<<<blocks_for(n, 256), 256>>>(n, idx, mval, larg, prm->sigma_curv, tvox);
                TN_HIP(hipGetLastError());
            }
        }
        if (prm->thick > 0.0f) {   // thin layers (src/v2m_grid.cpp:318-338)
            const float floor_mm = prm->thin_floor * vmin;
            tng::k_thin_size<<<nbv, 256>>>(nv, tvox, vmin, prm->thick, floor_mm, h);
            hmin = std::min(hmin, floor_mm);
        }
        mark("thickness");
        // the interface fields (src/v2m_grid.cpp:343-468)
        const bool tpm_raw = tpm && prm->sigma <= 0.0f;
        if (gray) {
            const int m = vol->nthr;
            const float* dI = tmp.upload(vol->gray, (size_t)nv);
            G->gTW.assign(vol->thresholds, vol->thresholds + m);
            G->gTW.resize(2 * (size_t)m, 1.0f);
            int32_t* cnt = ops.counts(0);
            int32_t* steps = tmp.alloc<int32_t>((size_t)3 * nv);
            for (int k = 0; k < m; ++k) {   // W_k = 4 x the median step across t_k
                TN_HIP(hipMemset(cnt, 0, sizeof(int32_t)));
                for (int a = 0; a < 3; ++a) {
                    tng::k_gray_steps<<<nbv, 256>>>(f, dI, vol->thresholds[k], a, cnt, steps);
                }
                int32_t ns = 0;
                TN_HIP(hipMemcpy(&ns, cnt, sizeof ns, hipMemcpyDeviceToHost));
                if (ns > 0) {
                    const auto kv = ops.sort(steps, nullptr, ns, 31);
                    int32_t bitsv = 0;
                    TN_HIP(hipMemcpy(&bitsv, kv.first + ns / 2, sizeof bitsv, hipMemcpyDeviceToHost));
                    float med;
                    std::memcpy(&med, &bitsv, 4);
                    G->gTW[m + k] = std::max(1e-12f, 4.0f * med);
                }
            }
            float* dTW = mem.upload(G->gTW.data(), G->gTW.size());
            G->dTW = dTW;
            G->gI = mem.upload(vol->gray, (size_t)nv);
            if (nslot > 0) {
                tng::k_gray_phi<<<blocks_for((int64_t)nslot * TN_SLOT, 256), 256>>>(f, dI, dTW, m, slot_brick, slot_label,
                                                                                  nslot, phi);
            }
            I.gm = m;
        } else if (tpm_raw) {
            if (nslot > 0) {
                tng::k_source_phi<<<blocks_for((int64_t)nslot * TN_SLOT, 256), 256>>>(f, src, slot_brick, slot_label, nslot,
                                                                                     phi);
            }
        } else {
            smooth(src, prm->sigma, Ri, phi);
        }
        TN_HIP(hipGetLastError());
        mark("interface");
        if (!gray && !tpm_raw && prm->sigma_thin > 0.0f) {   // the thickness-adaptive blend
            float* w = tmp.alloc<float>((size_t)nv);
            float* t2 = tmp.alloc<float>((size_t)nv);
            TN_HIP(hipMemcpy(w, tvox, sizeof(float) * nv, hipMemcpyDeviceToDevice));
            tng::Taps none;
            std::memset(&none, 0, sizeof none);
            for (int a = 0; a < 3; ++a) {
                tng::k_axis_pass<<<nbv, 256>>>(f, a, 2, 1, none, w, t2);
                std::swap(w, t2);
            }
            tng::k_smoothstep<<<nbv, 256>>>(nv, prm->thin_lo, std::max(1e-3f, prm->thin_hi - prm->thin_lo), w);
            const tng::Taps ker = gauss_taps(1.0f, 3);
            for (int a = 0; a < 3; ++a) {
                tng::k_axis_pass<<<nbv, 256>>>(f, a, 3, 0, ker, w, t2);
                std::swap(w, t2);
            }
            float* phi_s = tmp.alloc<float>(std::max<size_t>(1, nphi));
            if (nphi) {
                TN_HIP(hipMemcpy(phi_s, phi, sizeof(float) * nphi, hipMemcpyDeviceToDevice));
            }
            smooth(src, prm->sigma_thin, smooth_radius(prm->sigma_thin), phi);
            if (nslot > 0) {
                tng::k_blend<<<blocks_for((int64_t)nslot * TN_SLOT, 256), 256>>>(f, slot_brick, nslot, w, phi_s, phi);
            }
            TN_HIP(hipGetLastError());
        }
        if (!gray && !tpm && prm->preserve > 0.0f) {
            tng::k_preserve<<<nbv, 256>>>(f, prm->preserve, phi);
        }
        mark("blend");
        // 5. gradient limiting: Jacobi sweeps until nothing changes
        // (over the bricks that can still change: all at first, then those within
        // one brick of a change; the commit keeps h whole for the next sweep's reads)
        float* hn = tmp.alloc<float>((size_t)nv);
        int32_t* bchanged = tmp.alloc<int32_t>(nb);
        int32_t* list = tmp.alloc<int32_t>(nb);
        int32_t* list2 = tmp.alloc<int32_t>(nb);
        int32_t* cnt = ops.counts(0);
        k_iota<<<blocks_for(nb, 256), 256>>>(list, nb);
        int32_t nlist = nb;
        I.limit_sweeps = 0;
        while (nlist > 0) {
            TN_HIP(hipMemset(bchanged, 0, sizeof(int32_t) * nb));
            tng::k_limit_bricks<<<nlist, 512>>>(f, prm->g, list, h, hn, bchanged);
            tng::k_commit_bricks<<<nlist, 512>>>(f, list, hn, h);
            ++I.limit_sweeps;
            TN_HIP(hipMemset(cnt, 0, sizeof(int32_t)));
            tng::k_next_bricks<<<blocks_for(nb, 256), 256>>>(f, bchanged, cnt, list2);
            TN_HIP(hipGetLastError());
            TN_HIP(hipMemcpy(&nlist, cnt, sizeof nlist, hipMemcpyDeviceToHost));
            std::swap(list, list2);
        }
        f.h = h;
        mark("limit");
        // 6. grades
        G->grade = mem.alloc<uint8_t>((size_t)nv);
        if (hmax <= hmin * 1.0001f) {
            tng::k_fill_u8<<<nbv, 256>>>(G->grade, nv, 0);
        } else {
            const float lden = std::log(hmax / hmin);
            int32_t* cnt = ops.counts(0);
            int32_t* doubt = tmp.alloc<int32_t>((size_t)nv);
            tng::k_grade<<<nbv, 256>>>(nv, h, hmin, lden, cnt, doubt, G->grade);
            TN_HIP(hipGetLastError());
            int32_t n = 0;
            TN_HIP(hipMemcpy(&n, cnt, sizeof n, hipMemcpyDeviceToHost));
            if (n > 0) {   // the host's logf for the voxels whose bin was in doubt
                float* hv = tmp.alloc<float>((size_t)n);
                uint8_t* gv = tmp.alloc<uint8_t>((size_t)n);
                tng::k_gather_f32<<<blocks_for(n, 256), 256>>>(n, doubt, h, hv);
                std::vector<float> hh((size_t)n);
                TN_HIP(hipMemcpy(hh.data(), hv, sizeof(float) * n, hipMemcpyDeviceToHost));
                std::vector<uint8_t> gg((size_t)n);
                for (int32_t q = 0; q < n; ++q) {
                    float t = std::log(hh[q] / hmin) / lden;
                    t = t < 0.0f ? 0.0f : (t > 1.0f ? 1.0f : t);
                    gg[q] = (uint8_t)(t * 255.0f + 0.5f);
                }
                TN_HIP(hipMemcpy(gv, gg.data(), (size_t)n, hipMemcpyHostToDevice));
                tng::k_scatter_u8<<<blocks_for(n, 256), 256>>>(n, doubt, gv, G->grade);
                TN_HIP(hipGetLastError());
            }
        }
        mark("grade");
        {   // the largest h (the hash's pruning test), and the Field completed for seeding
            int32_t* mm = tmp.alloc<int32_t>(2);
            const int32_t init[2] = { 0x7F800000, 0 };
            TN_HIP(hipMemcpy(mm, init, sizeof init, hipMemcpyHostToDevice));
            tng::k_pos_minmax<<<nbv, 256>>>(h, nv, mm);
            int32_t got[2];
            TN_HIP(hipMemcpy(got, mm, sizeof got, hipMemcpyDeviceToHost));
            std::memcpy(&G->hbig, &got[1], 4);
            f.gm = I.gm;
            f.gI = G->gI;
            f.gTW = G->dTW;
            set_reciprocals(f);
            tnd::Field* self = mem.alloc<tnd::Field>(1);
            f.self = self;
            TN_HIP(hipMemcpy(self, &f, sizeof f, hipMemcpyHostToDevice));
        }
        TN_HIP(hipDeviceSynchronize());
        if (verbose && !marks.empty()) {
            std::fprintf(stderr, "[tn_grid_build] %d x %d x %d, %d slots, %d sweeps:", nx, ny, nz, nslot, I.limit_sweeps);
            for (size_t k = 1; k < marks.size(); ++k) {
                float ms = 0.0f;
                TN_HIP(hipEventElapsedTime(&ms, marks[k - 1].second, marks[k].second));
                std::fprintf(stderr, " %s %.2f", marks[k].first, ms);
            }
            std::fprintf(stderr, " (ms)\n");
            for (auto& m : marks) {
                (void)hipEventDestroy(m.second);
            }
        }
        I.hbase = hbase;
        I.hmin = hmin;
        I.hmax = hmax;
        I.hcurv = hcurv;
        *out = G.release();
    });
}

int tn_grid_get_info(const tn_grid* g, tn_grid_info* info) {
    return guarded([&] {
        arg_check(g && info, "null argument");
        *info = g->info;
    });
}

int tn_grid_download(const tn_grid* g, int32_t* bl_cnt, uint16_t* bl_lab, int32_t* bl_slot, float* phi, float* h,
                     uint8_t* grade, float* gTW) {
    return guarded([&] {
        arg_check(g != nullptr, "null grid");
        TN_HIP(hipSetDevice(g->mem.device));
        const tn_grid_info& I = g->info;
        const size_t nv = (size_t)I.nx * I.ny * I.nz, nb = (size_t)I.nbx * I.nby * I.nbz;
        auto get = [](void* dst, const void* src, size_t bytes) {
            if (dst && bytes) {
                TN_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
            }
        };
        get(bl_cnt, g->f.bl_cnt, nb * sizeof(int32_t));
        get(bl_lab, g->f.bl_lab, nb * TN_BL * sizeof(uint16_t));
        get(bl_slot, g->f.bl_slot, nb * sizeof(int32_t));
        get(phi, g->f.phi, (size_t)I.nslot * TN_SLOT * sizeof(float));
        get(h, g->f.h, nv * sizeof(float));
        get(grade, g->grade, nv);
        if (gTW && !g->gTW.empty()) {
            std::memcpy(gTW, g->gTW.data(), g->gTW.size() * sizeof(float));
        }
    });
}

int tn_grid_destroy(tn_grid* g) {
    return guarded([&] { delete g; });
}

int tn_seed_count_from_grid(const tn_grid* g, const tn_relax_params* prm, int64_t* n) {
    return guarded([&] {
        arg_check(g != nullptr && n != nullptr && prm != nullptr, "null argument");
        TN_HIP(hipSetDevice(g->mem.device));
        g->seeds.reset();
        DevMem mem;
        mem.device = g->mem.device;
        const Seeds s = seeding_core(g->f, g->grade, g->info.hmin, g->info.hmax, prm, mem);
        // keep the nodes (not the seeding's work arrays) for tn_seed_from_grid
        auto c = std::make_unique<tn_grid::SeedCache>();
        c->mem.device = g->mem.device;
        c->prm = *prm;
        c->s.n = s.n;
        auto keep = [&](auto* src, size_t count) {
            using T = std::remove_pointer_t<decltype(src)>;
            T* d = c->mem.alloc<T>(std::max<size_t>(count, 1));
            if (count && src) {
                TN_HIP(hipMemcpy(d, src, count * sizeof(T), hipMemcpyDeviceToDevice));
            }
            return src ? d : nullptr;
        };
        const size_t m = (size_t)s.n;
        c->s.P = keep(s.P, 3 * m);
        c->s.lab = keep(s.lab, m);
        c->s.typ = keep(s.typ, m);
        c->s.part = keep(s.part, 2 * m);
        c->s.part3 = keep(s.part3, m);
        g->seeds = std::move(c);
        *n = s.n;
    });
}

int tn_seed_from_grid(const tn_grid* g, const tn_relax_params* prm, int64_t n, float* P, uint16_t* lab, uint8_t* typ,
                      uint16_t* part, uint16_t* part3) {
    return guarded([&] {
        arg_check(g != nullptr, "null grid");
        arg_check(n >= 0 && (n == 0 || (P && lab && typ && part)), "seeding: null output array");
        arg_check(prm != nullptr, "null parameters");
        TN_HIP(hipSetDevice(g->mem.device));
        DevMem mem;
        mem.device = g->mem.device;
        // the nodes of the count call with these parameters, else a fresh seeding
        // (deterministic: the same nodes)
        std::unique_ptr<tn_grid::SeedCache> cached = std::move(g->seeds);
        const bool hit = cached && std::memcmp(&cached->prm, prm, sizeof *prm) == 0;
        const Seeds s = hit ? cached->s : seeding_core(g->f, g->grade, g->info.hmin, g->info.hmax, prm, mem);
        arg_check(s.n == n, "seeding: the output holds " + std::to_string(n) + " nodes, the seeding made " +
                                std::to_string(s.n) + " (tn_seed_count_from_grid)");
        if (s.n == 0) {
            return;
        }
        TN_HIP(hipMemcpy(P, s.P, sizeof(float) * 3 * s.n, hipMemcpyDeviceToHost));
        TN_HIP(hipMemcpy(lab, s.lab, sizeof(uint16_t) * s.n, hipMemcpyDeviceToHost));
        TN_HIP(hipMemcpy(typ, s.typ, s.n, hipMemcpyDeviceToHost));
        TN_HIP(hipMemcpy(part, s.part, sizeof(uint16_t) * 2 * s.n, hipMemcpyDeviceToHost));
        if (part3) {
            TN_HIP(hipMemcpy(part3, s.part3, sizeof(uint16_t) * s.n, hipMemcpyDeviceToHost));
        }
    });
}

int tn_plan_create_from_grid(const tn_grid* g, const tn_relax_params* prm, const tn_nodes_view* seeds,
                             tn_plan** out) {
    return guarded([&] {
        arg_check(g != nullptr && out != nullptr, "null argument");
        *out = nullptr;
        check_params(prm);
        check_nodes(seeds);
        const tn_grid_info& I = g->info;
        const float vs[3] = { g->f.vx, g->f.vy, g->f.vz };
        int64_t nkeys = 0;
        const tnd::Hash H = make_hash_geom(I.nx, I.ny, I.nz, vs, I.hmin, I.hmax, g->hbig, prm, &nkeys);
        // the plan keeps its own copy of the grid (device to device)
        *out = plan_build(prm, seeds, g->mem.device, H, nkeys, [&](DevMem& mem) {
            const size_t nv = (size_t)I.nx * I.ny * I.nz, nb = (size_t)I.nbx * I.nby * I.nbz;
            auto dup = [&](auto* src, size_t count) {
                using T = std::remove_const_t<std::remove_pointer_t<decltype(src)>>;
                T* d = mem.alloc<T>(std::max<size_t>(count, 1));
                if (count) {
                    TN_HIP(hipMemcpy(d, src, count * sizeof(T), hipMemcpyDeviceToDevice));
                }
                return d;
            };
            tnd::Field f = g->f;
            f.L = dup(g->f.L, nv);
            f.bl_cnt = dup(g->f.bl_cnt, nb);
            f.bl_lab = dup(g->f.bl_lab, nb * TN_BL);
            f.bl_slot = dup(g->f.bl_slot, nb);
            f.phi = dup(g->f.phi, (size_t)I.nslot * TN_SLOT);
            f.h = dup(g->f.h, nv);
            f.gI = I.gm > 0 ? dup(g->gI, nv) : nullptr;
            f.gTW = I.gm > 0 ? dup(g->dTW, 2 * (size_t)I.gm) : nullptr;
            tnd::Field* self = mem.alloc<tnd::Field>(1);
            f.self = self;
            TN_HIP(hipMemcpy(self, &f, sizeof f, hipMemcpyHostToDevice));
            return f;
        });
    });
}

}  // extern "C"
