// tn_adjoint.hpp -- a differentiable relaxation and its hand-written adjoint.
//
// No counterpart in the reference, whose relaxation is not differentiable: its bar
// law is a hinge (l0 - l)+ (src/opencl/v2m_particle_body.cl:161-184), its size is
// the nearest voxel's (v2m_h_at, :30-36), and its move has projections, type
// switches and FIRE resets. The surrogate here keeps the reference's force and
// Jacobi step (src/opencl/v2m_particle_body.cl:844-870, fire off) and smooths what
// is not differentiable:
//
//   h(x)       trilinear in a sizing grid theta with smoothstep cell fractions (C1 across
//              cell faces; gx x gy x gz, origin o, spacing d),
//              the position clamped to the grid
//   bar i->j   e = x_i - x_j, l = |e|, l0 = c_ij (h_i + h_j) / 2, u = l0 - l
//              (c_ij = fsurf if neither end is interior, else fscale)
//   force      F_i = sum_j sp(u) e / l,   sp(u) = log(1 + exp(beta u)) / beta
//   stiffness  D_i = 1 + sum_j s(beta u),  s = the logistic function (a smooth count
//              of the active bars, for max(#active, 1))
//   step       x_i <- x_i + m_i dt F_i / D_i  (m_i = 1 for interior nodes, else 0)
//
// and after T steps the loss
//
//   L = (1/n) sum_i sum_j sp(u_ij)^2 / 2  +  wq (1/ntet) sum_t (1 - q_t),
//   q_t = 12 (3 V_s)^(2/3) / sum_edges l^2,  V_s = sqrt(V^2 + eps_v^2)
//
// (Joe-Liu quality, src/v2m_tetra.cpp:100-120, over a fixed tetrahedral
// connectivity). The neighbour lists are fixed too. The adjoint runs the steps
// backwards from checkpoints of the positions every `ckpt_every` steps; within a
// segment the states are recomputed, with each node's F and D, and one gather
// pass per step forms the new adjoint from the node's own bars and the bars that
// point at it (a reverse list built once), so no adjoint sum needs an atomic.
// Only the sizing-grid gradient is scattered (double atomics).
#pragma once

namespace tnadj {

using tnk::kBlock;
using tnk::kWave;

struct AdjArgs {
    int n;
    const int32_t* __restrict__ nbr;     // [K * n] list position-major: entry k of node i at k * n + i
    const int32_t* __restrict__ nnb;     // [n]
    const uint8_t* __restrict__ typ;     // [n]
    const int32_t* __restrict__ rstart;  // [n + 1] reverse lists: the bars pointing at each node
    const int32_t* __restrict__ rsrc;    // [E] entries i * K + k
    const int32_t* __restrict__ tet;     // [4 ntet]
    const int32_t* __restrict__ vstart;  // [n + 1] the tets at each node
    const int32_t* __restrict__ vtet;    // entries 4 t + corner
    int ntet;
    const float* __restrict__ theta;
    const float4* __restrict__ B;        // [n] the start positions; states hold displacements from them
    int gx, gy, gz;
    float ox, oy, oz, dx, dy, dz;
    float dt, fscale, fsurf, beta, wq, eps_v;
};

// ---- the sizing grid ---------------------------------------------------------------

struct Tri {
    int i0, j0, k0;
    float fx, fy, fz;
    float gsx, gsy, gsz;   // d u / d x per axis (0 where the position is clamped)
};

__device__ __forceinline__ void tri_axis(float x, float o, float d, int g, int& i0, float& f, float& gs) {
    float u = (x - o) / d;
    gs = 1.0f / d;
    if (!(u > 0.0f)) {
        u = 0.0f;
        gs = 0.0f;
    } else if (u >= (float)(g - 1)) {
        u = (float)(g - 1);
        gs = 0.0f;
    }
    i0 = (int)floorf(u);
    if (i0 > g - 2) {
        i0 = g - 2;
    }
    // the cell fraction through smoothstep: the interpolant is C1 across cell faces, so
    // the loss has no kinks where nodes cross them
    const float t = u - (float)i0;
    f = t * t * (3.0f - 2.0f * t);
    gs *= 6.0f * t * (1.0f - t);
}

__device__ __forceinline__ Tri tri_at(const AdjArgs& A, float x, float y, float z) {
    Tri t;
    tri_axis(x, A.ox, A.dx, A.gx, t.i0, t.fx, t.gsx);
    tri_axis(y, A.oy, A.dy, A.gy, t.j0, t.fy, t.gsy);
    tri_axis(z, A.oz, A.dz, A.gz, t.k0, t.fz, t.gsz);
    return t;
}

__device__ __forceinline__ int tri_index(const AdjArgs& A, const Tri& t, int c) {
    return (t.i0 + (c & 1)) + A.gx * ((t.j0 + ((c >> 1) & 1)) + A.gy * (t.k0 + (c >> 2)));
}

__device__ __forceinline__ float tri_weight(const Tri& t, int c) {
    return ((c & 1) ? t.fx : 1.0f - t.fx) * (((c >> 1) & 1) ? t.fy : 1.0f - t.fy) * ((c >> 2) ? t.fz : 1.0f - t.fz);
}

__device__ __forceinline__ float size_of(const AdjArgs& A, float x, float y, float z) {
    const Tri t = tri_at(A, x, y, z);
    float h = 0.0f;
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        h += tri_weight(t, c) * A.theta[tri_index(A, t, c)];
    }
    return h;
}

// The grid gradient is accumulated in fixed point, 2^-44 units in 64-bit integers:
// integer additions give the same sum in any order, so the gradient is the same on
// every run (double atomics would round in the order the waves arrive). A term is
// rounded to the unit once; a cell's sum stays far inside the +-2^19 range.
constexpr double kThetaScale = 17592186044416.0;   // 2^44

// gtheta[idx] += v for every lane of the wave (all lanes call it): lanes of one
// run of equal cells (neighbouring nodes share cells) add once
__device__ __forceinline__ void theta_add(unsigned long long* __restrict__ gtheta, int idx, double v) {
    const int lane = threadIdx.x & (kWave - 1);
    const int prev = __shfl_up(idx, 1, kWave);
    const bool head = lane == 0 || prev != idx;
    const uint64_t heads = __ballot(head);
    const uint64_t upto = lane == kWave - 1 ? ~0ull : ((1ull << (lane + 1)) - 1);
    const int seg = __popcll(heads & upto);
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        const double y = __shfl_up(v, d, kWave);
        const int sy = __shfl_up(seg, d, kWave);
        if (lane >= d && sy == seg) {
            v += y;
        }
    }
    const int next = __shfl_down(seg, 1, kWave);
    if ((lane == kWave - 1 || next != seg) && v != 0.0) {
        atomicAdd(&gtheta[idx], (unsigned long long)__double2ll_rn(v * kThetaScale));
    }
}

// the adjoint hd of h at x: hd * dh/dx into g, hd * w_c into the grid gradient
// (every lane of the wave calls it; hd = 0 for a lane with nothing to add)
__device__ __forceinline__ void size_adjoint(const AdjArgs& A, float x, float y, float z, float hd, float3& g,
                                             unsigned long long* __restrict__ gtheta) {
    const Tri t = tri_at(A, x, y, z);
    float th[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        th[c] = A.theta[tri_index(A, t, c)];
        theta_add(gtheta, tri_index(A, t, c), (double)(hd * tri_weight(t, c)));
    }
    // d h / d f per axis (trilinear in the smoothed fractions), times d f / d x
    float hx = 0.0f, hy = 0.0f, hz = 0.0f;
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        const float wx = (c & 1) ? t.fx : 1.0f - t.fx, wy = ((c >> 1) & 1) ? t.fy : 1.0f - t.fy,
                    wz = (c >> 2) ? t.fz : 1.0f - t.fz;
        const float sx = (c & 1) ? 1.0f : -1.0f, sy = ((c >> 1) & 1) ? 1.0f : -1.0f, sz = (c >> 2) ? 1.0f : -1.0f;
        hx += sx * wy * wz * th[c];
        hy += wx * sy * wz * th[c];
        hz += wx * wy * sz * th[c];
    }
    g.x += hd * hx * t.gsx;
    g.y += hd * hy * t.gsy;
    g.z += hd * hz * t.gsz;
}

// ---- one bar -------------------------------------------------------------------------

__device__ __forceinline__ float softplus(float u, float beta) {
    return u > 0.0f ? u + log1pf(expf(-beta * u)) / beta : log1pf(expf(beta * u)) / beta;
}

__device__ __forceinline__ float logistic(float v) {
    return 1.0f / (1.0f + expf(-v));
}

__device__ __forceinline__ float bar_c(const AdjArgs& A, int ti, int tj) {
    return (ti != TN_INTERIOR && tj != TN_INTERIOR) ? A.fsurf : A.fscale;
}

struct Bar {
    float ex, ey, ez, l, sp, sg;   // e, |e|, sp(u), s(beta u)
    bool ok;
};

// bar i -> j from the start positions b and the states x (displacement, h): the
// start positions' difference is exact for neighbours (Sterbenz), so e keeps the
// displacements' precision however far the volume is from the origin
__device__ __forceinline__ Bar bar_at(const AdjArgs& A, float4 bi, float4 xi, float4 bj, float4 xj, float c) {
    Bar b;
    b.ex = (bi.x - bj.x) + (xi.x - xj.x);
    b.ey = (bi.y - bj.y) + (xi.y - xj.y);
    b.ez = (bi.z - bj.z) + (xi.z - xj.z);
    b.l = sqrtf(b.ex * b.ex + b.ey * b.ey + b.ez * b.ez);
    b.ok = b.l > 0.0f;
    const float u = c * 0.5f * (xi.w + xj.w) - b.l;
    b.sp = softplus(u, A.beta);
    b.sg = logistic(A.beta * u);
    return b;
}

// F_i and D_i of node i at the state X (xyz, h). The bar forces of a relaxed node
// nearly cancel, and the adjoint weighs F_i itself (kappa = mu . F / D^2): the sum
// runs in double so F_i keeps its relative precision.
__device__ __forceinline__ float4 force_and_stiffness(const AdjArgs& A, const float4* __restrict__ X, int i) {
    const float4 xi = X[i], bi = A.B[i];
    const int ti = A.typ[i], m = A.nnb[i];
    double Fx = 0.0, Fy = 0.0, Fz = 0.0, a = 0.0;
    for (int k = 0; k < m; ++k) {
        const int j = A.nbr[(size_t)k * A.n + i];
        const Bar b = bar_at(A, bi, xi, A.B[j], X[j], bar_c(A, ti, A.typ[j]));
        if (b.ok) {
            const double w = (double)b.sp / (double)b.l;
            Fx += w * b.ex;
            Fy += w * b.ey;
            Fz += w * b.ez;
            a += b.sg;
        }
    }
    return make_float4((float)Fx, (float)Fy, (float)Fz, (float)(1.0 + a));
}

// ---- forward ---------------------------------------------------------------------------

__global__ __launch_bounds__(kBlock) void k_adj_load(AdjArgs A, const float* __restrict__ x0, float4* __restrict__ B,
                                                     float4* __restrict__ X) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < A.n) {
        const float x = x0[3 * (size_t)i], y = x0[3 * (size_t)i + 1], z = x0[3 * (size_t)i + 2];
        B[i] = make_float4(x, y, z, 0.0f);
        X[i] = make_float4(0.0f, 0.0f, 0.0f, size_of(A, x, y, z));
    }
}

// X -> Xn (positions and their sizes); FD (optional): F and D of each node at X
__global__ __launch_bounds__(kBlock) void k_adj_step(AdjArgs A, const float4* __restrict__ X, float4* __restrict__ Xn,
                                                     float4* __restrict__ FD) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= A.n) {
        return;
    }
    const float4 fd = force_and_stiffness(A, X, i);
    if (FD) {
        FD[i] = fd;
    }
    float4 x = X[i];
    if (A.typ[i] == TN_INTERIOR) {
        const float s = A.dt / fd.w;
        x.x += s * fd.x;
        x.y += s * fd.y;
        x.z += s * fd.z;
        const float4 b = A.B[i];
        x.w = size_of(A, b.x + x.x, b.y + x.y, b.z + x.z);
    }
    Xn[i] = x;
}

// out = B + X (positions), or X alone (B null: an adjoint), 3 floats per node
__global__ __launch_bounds__(kBlock) void k_adj_unload(int n, const float4* __restrict__ B,
                                                       const float4* __restrict__ X, float* __restrict__ out) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) {
        float4 x = X[i];
        if (B) {
            const float4 b = B[i];
            x.x += b.x;
            x.y += b.y;
            x.z += b.z;
        }
        out[3 * (size_t)i] = x.x;
        out[3 * (size_t)i + 1] = x.y;
        out[3 * (size_t)i + 2] = x.z;
    }
}

// ---- the loss ---------------------------------------------------------------------------

// The tets' quality runs in double precision: near-flat tets (common in a
// Delaunay of lattice points) have a volume that cancels to a few ulps of its
// terms, and the quality's derivative scales with V / Vs^2.
struct D3 {
    double x, y, z;
};
__device__ __forceinline__ D3 d3(float4 b, float4 x) {
    return D3{ (double)b.x + (double)x.x, (double)b.y + (double)x.y, (double)b.z + (double)x.z };
}
__device__ __forceinline__ D3 sub3(D3 a, D3 b) {
    return D3{ a.x - b.x, a.y - b.y, a.z - b.z };
}
__device__ __forceinline__ D3 cross3(D3 a, D3 b) {
    return D3{ a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x };
}
__device__ __forceinline__ double dot3(D3 a, D3 b) {
    return a.x * b.x + a.y * b.y + a.z * b.z;
}

struct TetQ {
    double q, V, Vs, S;
    D3 a, b, c;   // x1 - x0, x2 - x0, x3 - x0
};

__device__ __forceinline__ TetQ tet_quality(const AdjArgs& A, const float4* __restrict__ X, int t) {
    const int4 v = *reinterpret_cast<const int4*>(A.tet + 4 * (size_t)t);
    const D3 p0 = d3(A.B[v.x], X[v.x]), p1 = d3(A.B[v.y], X[v.y]), p2 = d3(A.B[v.z], X[v.z]),
             p3 = d3(A.B[v.w], X[v.w]);
    TetQ r;
    r.a = sub3(p1, p0);
    r.b = sub3(p2, p0);
    r.c = sub3(p3, p0);
    r.V = dot3(r.a, cross3(r.b, r.c)) / 6.0;
    r.Vs = sqrt(r.V * r.V + (double)A.eps_v * (double)A.eps_v);
    const D3 d12 = sub3(p2, p1), d13 = sub3(p3, p1), d23 = sub3(p3, p2);
    r.S = dot3(r.a, r.a) + dot3(r.b, r.b) + dot3(r.c, r.c) + dot3(d12, d12) + dot3(d13, d13) + dot3(d23, d23);
    r.q = 12.0 * cbrt(9.0 * r.Vs * r.Vs) / r.S;
    return r;
}

// per block: (energy, 1 - q) partial sums
__global__ __launch_bounds__(kBlock) void k_adj_loss_parts(AdjArgs A, const float4* __restrict__ X,
                                                           double2* __restrict__ parts) {
    __shared__ double se[kBlock / kWave], sq[kBlock / kWave];
    double e = 0.0, q = 0.0;
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < A.n; i += gridDim.x * kBlock) {
        const float4 xi = X[i], bi = A.B[i];
        const int ti = A.typ[i], m = A.nnb[i];
        float s = 0.0f;
        for (int k = 0; k < m; ++k) {
            const int j = A.nbr[(size_t)k * A.n + i];
            const Bar b = bar_at(A, bi, xi, A.B[j], X[j], bar_c(A, ti, A.typ[j]));
            if (b.ok) {
                s += 0.5f * b.sp * b.sp;
            }
        }
        e += s;
    }
    for (int t = blockIdx.x * kBlock + threadIdx.x; t < A.ntet; t += gridDim.x * kBlock) {
        q += 1.0 - tet_quality(A, X, t).q;
    }
    for (int d = kWave / 2; d > 0; d >>= 1) {
        e += __shfl_xor(e, d, kWave);
        q += __shfl_xor(q, d, kWave);
    }
    const int w = threadIdx.x / kWave;
    if ((threadIdx.x & (kWave - 1)) == 0) {
        se[w] = e;
        sq[w] = q;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = 0.0, b = 0.0;
        for (int k = 0; k < kBlock / kWave; ++k) {
            a += se[k];
            b += sq[k];
        }
        parts[blockIdx.x] = make_double2(a, b);
    }
}

// one block: the loss from the partials (a fixed reduction tree)
__global__ __launch_bounds__(kBlock) void k_adj_loss_sum(AdjArgs A, const double2* __restrict__ parts, int m,
                                                         double* __restrict__ out) {
    __shared__ double se[kBlock], sq[kBlock];
    double e = 0.0, q = 0.0;
    for (int k = threadIdx.x; k < m; k += kBlock) {
        e += parts[k].x;
        q += parts[k].y;
    }
    se[threadIdx.x] = e;
    sq[threadIdx.x] = q;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            se[threadIdx.x] += se[threadIdx.x + w];
            sq[threadIdx.x] += sq[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double E = A.n > 0 ? se[0] / A.n : 0.0, Q = A.ntet > 0 ? sq[0] / A.ntet : 0.0;
        out[0] = E + (double)A.wq * Q;
        out[1] = E;
        out[2] = Q;
    }
}

// d L / d x_T and the grid gradient of the loss: the energy of the node's own bars
// and of the bars pointing at it, and the quality of its tets
__global__ __launch_bounds__(kBlock) void k_adj_loss_grad(AdjArgs A, const float4* __restrict__ X,
                                                          float4* __restrict__ lam, unsigned long long* __restrict__ gtheta) {
    int p = blockIdx.x * kBlock + threadIdx.x;
    const bool valid = p < A.n;   // the others stand in for node n - 1: every lane joins the scatter
    if (!valid) {
        p = A.n - 1;
    }
    const float en = 1.0f / (float)A.n;
    const float4 xp = X[p], bp = A.B[p];
    const int tp = A.typ[p];
    float3 g = make_float3(0.0f, 0.0f, 0.0f);
    float hd = 0.0f;
    // own bars p -> j: dE/de = -(1/n) sp s n, dE/dl0 = (1/n) sp s
    const int m = A.nnb[p];
    for (int k = 0; k < m; ++k) {
        const int j = A.nbr[(size_t)k * A.n + p];
        const float c = bar_c(A, tp, A.typ[j]);
        const Bar b = bar_at(A, bp, xp, A.B[j], X[j], c);
        if (b.ok) {
            const float w = en * b.sp * b.sg;
            g.x -= w * b.ex / b.l;
            g.y -= w * b.ey / b.l;
            g.z -= w * b.ez / b.l;
            hd += 0.5f * c * w;
        }
    }
    // bars i -> p
    for (int r = A.rstart[p]; r < A.rstart[p + 1]; ++r) {
        const int i = A.rsrc[r] / TN_K;
        const float c = bar_c(A, A.typ[i], tp);
        const Bar b = bar_at(A, A.B[i], X[i], bp, xp, c);
        if (b.ok) {
            const float w = en * b.sp * b.sg;
            g.x += w * b.ex / b.l;
            g.y += w * b.ey / b.l;
            g.z += w * b.ez / b.l;
            hd += 0.5f * c * w;
        }
    }
    // tets: d(-wq/ntet q)/dx_p, q = 12 (3 Vs)^(2/3) / S (in double)
    const double wt = A.ntet > 0 ? -(double)A.wq / (double)A.ntet : 0.0;
    double tx = 0.0, ty = 0.0, tz = 0.0;
    const D3 pp = d3(bp, xp);
    for (int r = A.vstart[p]; r < A.vstart[p + 1]; ++r) {
        const int t = A.vtet[r] >> 2, corner = A.vtet[r] & 3;
        const TetQ Q = tet_quality(A, X, t);
        // d(6V)/dx_corner
        D3 dv;
        if (corner == 1) {
            dv = cross3(Q.b, Q.c);
        } else if (corner == 2) {
            dv = cross3(Q.c, Q.a);
        } else if (corner == 3) {
            dv = cross3(Q.a, Q.b);
        } else {
            const D3 u = cross3(Q.b, Q.c), v = cross3(Q.c, Q.a), w = cross3(Q.a, Q.b);
            dv = D3{ -(u.x + v.x + w.x), -(u.y + v.y + w.y), -(u.z + v.z + w.z) };
        }
        // d S / dx_corner = 2 sum over the other corners of (x_corner - x_other)
        const int4 v4 = *reinterpret_cast<const int4*>(A.tet + 4 * (size_t)t);
        const int vv[4] = { v4.x, v4.y, v4.z, v4.w };
        D3 ds{ 0.0, 0.0, 0.0 };
        for (int o = 0; o < 4; ++o) {
            if (o != corner) {
                const D3 xo = d3(A.B[vv[o]], X[vv[o]]);
                ds.x += 2.0 * (pp.x - xo.x);
                ds.y += 2.0 * (pp.y - xo.y);
                ds.z += 2.0 * (pp.z - xo.z);
            }
        }
        // dq = q ((2/3) dVs / Vs - dS / S), dVs = (V / Vs) dV, dV = dv / 6
        const double kv = Q.q * (2.0 / 3.0) * (Q.V / (Q.Vs * Q.Vs)) / 6.0, ks = Q.q / Q.S;
        tx += wt * (kv * dv.x - ks * ds.x);
        ty += wt * (kv * dv.y - ks * ds.y);
        tz += wt * (kv * dv.z - ks * ds.z);
    }
    g.x += (float)tx;
    g.y += (float)ty;
    g.z += (float)tz;
    size_adjoint(A, bp.x + xp.x, bp.y + xp.y, bp.z + xp.z, valid ? hd : 0.0f, g, gtheta);
    if (valid) {
        lam[p] = make_float4(g.x, g.y, g.z, 0.0f);
    }
}

// ---- one step backwards ------------------------------------------------------------------

// the adjoint of bar i -> j for node i's (nu, kappa): returns g_e (to x_i; -g_e to
// x_j) and g_l0 in .w
__device__ __forceinline__ float4 bar_adjoint(const Bar& b, float3 nu, float kappa, float beta) {
    const float inv = 1.0f / b.l;
    const float nx = b.ex * inv, ny = b.ey * inv, nz = b.ez * inv;
    const float vn = nu.x * nx + nu.y * ny + nu.z * nz;
    const float rho = beta * b.sg * (1.0f - b.sg);
    const float a = b.sp * inv, r = -b.sg * vn - a * vn + kappa * rho;
    return make_float4(a * nu.x + r * nx, a * nu.y + r * ny, a * nu.z + r * nz, b.sg * vn - kappa * rho);
}

// (nu, kappa) of node i: mu = m_i dt lambda'_i, nu = mu / D, kappa = mu . F / D^2
__device__ __forceinline__ float4 nu_kappa(const AdjArgs& A, int ti, float4 lamn, float4 fd) {
    if (ti != TN_INTERIOR) {
        return make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
    const float s = A.dt / fd.w;
    const float3 nu = make_float3(s * lamn.x, s * lamn.y, s * lamn.z);
    return make_float4(nu.x, nu.y, nu.z, (nu.x * fd.x + nu.y * fd.y + nu.z * fd.z) / fd.w);
}

__global__ __launch_bounds__(kBlock) void k_adj_nu(AdjArgs A, const float4* __restrict__ FD,
                                                   const float4* __restrict__ lamn, float4* __restrict__ nk) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i < A.n) {
        nk[i] = nu_kappa(A, A.typ[i], lamn[i], FD[i]);
    }
}

// lambda_k from lambda_{k+1} at the state X_k (nk: each node's (nu, kappa) at X_k)
// With FDprev, also (nu, kappa) of each node for the step before (k - 1), from the
// lambda_k just formed and that step's F, D: the next launch needs no separate pass.
__global__ __launch_bounds__(kBlock) void k_adj_back(AdjArgs A, const float4* __restrict__ X,
                                                     const float4* __restrict__ nk, const float4* __restrict__ lamn,
                                                     float4* __restrict__ lam, unsigned long long* __restrict__ gtheta,
                                                     const float4* __restrict__ FDprev, float4* __restrict__ nk_prev) {
    int p = blockIdx.x * kBlock + threadIdx.x;
    const bool valid = p < A.n;   // the others stand in for node n - 1: every lane joins the scatter
    if (!valid) {
        p = A.n - 1;
    }
    const float4 xp = X[p], bp = A.B[p];
    const int tp = A.typ[p];
    const float4 lp = lamn[p];
    float3 g = make_float3(lp.x, lp.y, lp.z);
    float hd = 0.0f;
    if (tp == TN_INTERIOR) {   // own bars p -> j
        const float4 nkp = nk[p];
        const float3 nu = make_float3(nkp.x, nkp.y, nkp.z);
        const int m = A.nnb[p];
        for (int k = 0; k < m; ++k) {
            const int j = A.nbr[(size_t)k * A.n + p];
            const float c = bar_c(A, tp, A.typ[j]);
            const Bar b = bar_at(A, bp, xp, A.B[j], X[j], c);
            if (b.ok) {
                const float4 ge = bar_adjoint(b, nu, nkp.w, A.beta);
                g.x += ge.x;
                g.y += ge.y;
                g.z += ge.z;
                hd += 0.5f * c * ge.w;
            }
        }
    }
    for (int r = A.rstart[p]; r < A.rstart[p + 1]; ++r) {   // bars i -> p of moving nodes i
        const int i = A.rsrc[r] / TN_K;
        const int ti = A.typ[i];
        if (ti != TN_INTERIOR) {
            continue;
        }
        const float4 nki = nk[i];
        const float c = bar_c(A, ti, tp);
        const Bar b = bar_at(A, A.B[i], X[i], bp, xp, c);
        if (b.ok) {
            const float4 ge = bar_adjoint(b, make_float3(nki.x, nki.y, nki.z), nki.w, A.beta);
            g.x -= ge.x;
            g.y -= ge.y;
            g.z -= ge.z;
            hd += 0.5f * c * ge.w;
        }
    }
    size_adjoint(A, bp.x + xp.x, bp.y + xp.y, bp.z + xp.z, valid ? hd : 0.0f, g, gtheta);
    if (valid) {
        const float4 l = make_float4(g.x, g.y, g.z, 0.0f);
        lam[p] = l;
        if (FDprev) {
            nk_prev[p] = nu_kappa(A, tp, l, FDprev[p]);
        }
    }
}

__global__ __launch_bounds__(kBlock) void k_adj_theta_out(int m, const unsigned long long* __restrict__ g,
                                                          float* __restrict__ out) {
    const int c = blockIdx.x * kBlock + threadIdx.x;
    if (c < m) {
        out[c] = (float)((double)(long long)g[c] * (1.0 / kThetaScale));
    }
}

__global__ __launch_bounds__(kBlock) void k_fill_f64(double* __restrict__ a, int64_t m) {
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < m) {
        a[i] = 0.0;
    }
}

// reverse-list keys: the target of every valid list entry (-1: none)
__global__ __launch_bounds__(kBlock) void k_adj_targets(int n, const int32_t* __restrict__ nbr,
                                                        const int32_t* __restrict__ nnb, int32_t* __restrict__ key,
                                                        int32_t* __restrict__ cnt) {
    const int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (e >= (int64_t)n * TN_K) {
        return;
    }
    const int i = (int)(e / TN_K), k = (int)(e % TN_K);
    if (k < nnb[i]) {
        const int j = nbr[(size_t)k * n + i];
        key[e] = j;
        atomicAdd(&cnt[j], 1);
    } else {
        key[e] = n;   // sorts past every node
    }
}

__global__ __launch_bounds__(kBlock) void k_adj_tet_keys(int ntet, const int32_t* __restrict__ tet,
                                                         int32_t* __restrict__ cnt) {
    const int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (e < 4 * (int64_t)ntet) {
        atomicAdd(&cnt[tet[e]], 1);
    }
}

}  // namespace tnadj

// ---- host side (included into tn_relax.hip, after its helpers) ----------------------------

struct tn_adj {
    int device = 0;
    tn_adj_problem prm{};
    int n = 0, ntet = 0, gcells = 0, iters = 0, ckpt = 1, nseg = 0, nparts = 1;
    DevMem mem;
    tnadj::AdjArgs A{};
    float4* ckpts = nullptr;   // [nseg * n]: the state at the start of each segment
    // two segment buffers, [(ckpt + 1) * n] states after a segment's first (which is its
    // checkpoint) and [ckpt * n] F and D: the backward of one segment runs beside the
    // recomputation of the one before it, on a second stream
    float4* seg[2] = { nullptr, nullptr };
    float4* fd[2] = { nullptr, nullptr };
    hipStream_t side = nullptr;
    hipEvent_t ev_start = nullptr, ev_ready[2] = { nullptr, nullptr }, ev_free[2] = { nullptr, nullptr };
    // ckpt_every == 1: every state is a checkpoint and the forward keeps F, D of each
    // ([iters * n]), so the backward recomputes nothing
    float4* fd_all = nullptr;
    float4* last = nullptr;    // the final state of the last forward
    float4 *Xa = nullptr, *Xb = nullptr, *lamA = nullptr, *lamB = nullptr, *nk = nullptr, *nk2 = nullptr,
           *base = nullptr;
    unsigned long long* gtheta = nullptr;   // fixed point (kThetaScale)
    double2* parts = nullptr;
    double* dloss = nullptr;
    bool forward_done = false;
    double loss[3] = { 0, 0, 0 };
    ~tn_adj() {
        for (hipEvent_t e : { ev_start, ev_ready[0], ev_ready[1], ev_free[0], ev_free[1] }) {
            if (e) {
                (void)hipEventDestroy(e);
            }
        }
        if (side) {
            (void)hipStreamDestroy(side);
        }
    }
};

namespace {

void adj_check(const tn_adj_problem* p) {
    arg_check(p != nullptr, "adjoint: null problem");
    arg_check(p->n > 0 && p->n < ((int64_t)1 << 31) / TN_K, "adjoint: node count out of range");
    arg_check(p->nbr && p->nnb && p->typ, "adjoint: null lists or types");
    arg_check(p->ntet >= 0 && p->ntet < ((int64_t)1 << 29) && (p->ntet == 0 || p->tet), "adjoint: bad tets");
    arg_check(p->gx >= 2 && p->gy >= 2 && p->gz >= 2, "adjoint: the sizing grid needs 2+ points per axis");
    arg_check(p->spacing[0] > 0 && p->spacing[1] > 0 && p->spacing[2] > 0, "adjoint: grid spacing must be positive");
    arg_check(p->iters >= 0 && p->ckpt_every >= 1, "adjoint: iters >= 0 and ckpt_every >= 1");
    arg_check(p->beta > 0 && p->dt > 0, "adjoint: beta and dt must be positive");
    for (int64_t i = 0; i < p->n; ++i) {
        arg_check(p->nnb[i] >= 0 && p->nnb[i] <= TN_K, "adjoint: list length out of range");
        for (int k = 0; k < p->nnb[i]; ++k) {
            const int32_t j = p->nbr[i * TN_K + k];
            arg_check(j >= 0 && j < p->n, "adjoint: list entry out of range");
        }
    }
    for (int64_t e = 0; e < 4 * p->ntet; ++e) {
        arg_check(p->tet[e] >= 0 && p->tet[e] < p->n, "adjoint: tet vertex out of range");
    }
}

// stable CSR of `m` keys in [0, n] (n: no entry): starts[n + 1] and the entries'
// indices in key order
void adj_csr(SeedOps& ops, int32_t* keys, int64_t m, int n, const int32_t* cnt_in, int32_t*& starts,
             int32_t*& entries) {
    starts = const_cast<int32_t*>(cnt_in);
    ops.scan(starts, n);
    if (m == 0) {
        entries = ops.mem.alloc<int32_t>(1);
        return;
    }
    const auto r = ops.sort(keys, nullptr, (int)m, bit_length((uint32_t)n));
    entries = r.second;
}

}  // namespace

extern "C" {

int tn_adj_create(const tn_adj_problem* p, int32_t device, tn_adj** out) {
    return guarded([&] {
        arg_check(out != nullptr, "adjoint: null handle");
        *out = nullptr;
        adj_check(p);
        std::unique_ptr<tn_adj> a(new tn_adj);
        a->device = pick_device(device);
        TN_HIP(hipSetDevice(a->device));
        a->mem.device = a->device;
        a->prm = *p;
        a->n = (int)p->n;
        a->ntet = (int)p->ntet;
        a->gcells = p->gx * p->gy * p->gz;
        a->iters = p->iters;
        a->ckpt = p->ckpt_every;
        a->nseg = (p->iters + a->ckpt - 1) / a->ckpt;
        const int n = a->n;
        DevMem& mem = a->mem;
        SeedOps ops{ mem };
        tnadj::AdjArgs& A = a->A;
        A.n = n;
        {   // the lists, list position-major (coalesced per position across nodes)
            std::vector<int32_t> t((size_t)n * TN_K, 0);
            for (int i = 0; i < n; ++i) {
                for (int k = 0; k < p->nnb[i]; ++k) {
                    t[(size_t)k * n + i] = p->nbr[(size_t)i * TN_K + k];
                }
            }
            A.nbr = mem.upload(t.data(), t.size());
        }
        A.nnb = mem.upload(p->nnb, (size_t)n);
        A.typ = mem.upload(p->typ, (size_t)n);
        A.ntet = a->ntet;
        A.tet = a->ntet ? mem.upload(p->tet, (size_t)4 * a->ntet) : mem.alloc<int32_t>(4);
        A.gx = p->gx;
        A.gy = p->gy;
        A.gz = p->gz;
        A.ox = p->origin[0];
        A.oy = p->origin[1];
        A.oz = p->origin[2];
        A.dx = p->spacing[0];
        A.dy = p->spacing[1];
        A.dz = p->spacing[2];
        A.dt = p->dt;
        A.fscale = p->fscale;
        A.fsurf = p->fsurf;
        A.beta = p->beta;
        A.wq = p->wq;
        A.eps_v = p->eps_v;
        // the bars pointing at each node, in (source, list position) order
        {
            const int64_t m = (int64_t)n * TN_K;
            int32_t* key = mem.alloc<int32_t>((size_t)m);
            int32_t* cnt = ops.counts(n);
            tnadj::k_adj_targets<<<blocks_for(m, tnk::kBlock), tnk::kBlock>>>(n, A.nbr, A.nnb, key, cnt);
            int32_t *starts = nullptr, *entries = nullptr;
            adj_csr(ops, key, m, n, cnt, starts, entries);
            A.rstart = starts;
            A.rsrc = entries;
        }
        // the tets at each node, in (tet, corner) order
        {
            const int64_t m = 4 * (int64_t)a->ntet;
            int32_t* cnt = ops.counts(n);
            int32_t* key = mem.alloc<int32_t>((size_t)std::max<int64_t>(m, 1));
            if (m) {
                TN_HIP(hipMemcpy(key, A.tet, sizeof(int32_t) * m, hipMemcpyDeviceToDevice));
                tnadj::k_adj_tet_keys<<<blocks_for(m, tnk::kBlock), tnk::kBlock>>>(a->ntet, A.tet, cnt);
            }
            int32_t *starts = nullptr, *entries = nullptr;
            adj_csr(ops, key, m, n, cnt, starts, entries);
            A.vstart = starts;
            A.vtet = entries;
        }
        a->ckpts = mem.alloc<float4>((size_t)std::max(a->nseg, 1) * n);
        if (a->ckpt == 1 && a->iters > 0) {
            a->fd_all = mem.alloc<float4>((size_t)a->iters * n);
        } else if (a->iters > 0) {
            for (int b = 0; b < (a->nseg > 1 ? 2 : 1); ++b) {
                a->seg[b] = mem.alloc<float4>((size_t)(a->ckpt + 1) * n);
                a->fd[b] = mem.alloc<float4>((size_t)a->ckpt * n);
            }
        }
        a->Xa = mem.alloc<float4>(n);
        a->Xb = mem.alloc<float4>(n);
        a->lamA = mem.alloc<float4>(n);
        a->lamB = mem.alloc<float4>(n);
        a->nk = mem.alloc<float4>(n);
        a->nk2 = mem.alloc<float4>(n);
        a->base = mem.alloc<float4>(n);
        A.B = a->base;
        a->gtheta = mem.alloc<unsigned long long>(a->gcells);
        a->nparts = (int)std::min<unsigned>(2048, blocks_for(std::max(n, a->ntet), tnk::kBlock));
        a->parts = mem.alloc<double2>(a->nparts);
        a->dloss = mem.alloc<double>(3);
        TN_HIP(hipGetLastError());
        TN_HIP(hipDeviceSynchronize());
        *out = a.release();
    });
}

int tn_adj_forward(tn_adj* a, const float* x0, const float* theta, float* xT, double* loss, void* stream) {
    return guarded([&] {
        arg_check(a != nullptr && x0 != nullptr && theta != nullptr, "adjoint forward: null argument");
        TN_HIP(hipSetDevice(a->device));
        const hipStream_t s = static_cast<hipStream_t>(stream);
        const int n = a->n;
        const unsigned nb = blocks_for(n, tnk::kBlock);
        tnadj::AdjArgs A = a->A;
        A.theta = theta;
        // state k lives in ckpts[k / ckpt] when k is a segment start, else in Xa / Xb
        float4* cur = a->ckpts;
        tnadj::k_adj_load<<<nb, tnk::kBlock, 0, s>>>(A, x0, a->base, cur);
        for (int k = 0; k < a->iters; ++k) {
            const int k1 = k + 1;
            float4* nxt = (k1 % a->ckpt == 0 && k1 < a->iters) ? a->ckpts + (size_t)(k1 / a->ckpt) * n
                                                               : (cur == a->Xa ? a->Xb : a->Xa);
            tnadj::k_adj_step<<<nb, tnk::kBlock, 0, s>>>(A, cur, nxt, a->fd_all ? a->fd_all + (size_t)k * n : nullptr);
            cur = nxt;
        }
        a->last = cur;
        tnadj::k_adj_loss_parts<<<a->nparts, tnk::kBlock, 0, s>>>(A, cur, a->parts);
        tnadj::k_adj_loss_sum<<<1, tnk::kBlock, 0, s>>>(A, a->parts, a->nparts, a->dloss);
        if (xT) {
            tnadj::k_adj_unload<<<nb, tnk::kBlock, 0, s>>>(n, a->base, cur, xT);
        }
        TN_HIP(hipGetLastError());
        TN_HIP(hipMemcpyAsync(a->loss, a->dloss, sizeof a->loss, hipMemcpyDeviceToHost, s));
        TN_HIP(hipStreamSynchronize(s));
        if (loss) {
            loss[0] = a->loss[0];
        }
        a->forward_done = true;
    });
}

int tn_adj_backward(tn_adj* a, const float* theta, float* grad_x0, float* grad_theta, void* stream) {
    return guarded([&] {
        arg_check(a != nullptr && theta != nullptr, "adjoint backward: null argument");
        arg_check(a->forward_done, "adjoint backward: no forward pass on this handle");
        TN_HIP(hipSetDevice(a->device));
        const hipStream_t s = static_cast<hipStream_t>(stream);
        const int n = a->n;
        const unsigned nb = blocks_for(n, tnk::kBlock);
        tnadj::AdjArgs A = a->A;
        A.theta = theta;
        tnadj::k_fill_f64<<<blocks_for(a->gcells, tnk::kBlock), tnk::kBlock, 0, s>>>(
            reinterpret_cast<double*>(a->gtheta), a->gcells);   // (all-zero bits)
        // the final state: recompute from the last checkpoint
        const int T = a->iters;
        float4* lam = a->lamA;
        float4* lamo = a->lamB;
        if (T == 0) {
            tnadj::k_adj_loss_grad<<<nb, tnk::kBlock, 0, s>>>(A, a->ckpts, lam, a->gtheta);
        }
        if (a->fd_all) {   // every state and its F, D kept by the forward
            tnadj::k_adj_loss_grad<<<nb, tnk::kBlock, 0, s>>>(A, a->last, lam, a->gtheta);
            float4 *nkc = a->nk, *nkn = a->nk2;
            if (T > 0) {
                tnadj::k_adj_nu<<<nb, tnk::kBlock, 0, s>>>(A, a->fd_all + (size_t)(T - 1) * n, lam, nkc);
            }
            for (int k = T - 1; k >= 0; --k) {   // each step also forms the next one's (nu, kappa)
                tnadj::k_adj_back<<<nb, tnk::kBlock, 0, s>>>(A, a->ckpts + (size_t)k * n, nkc, lam, lamo, a->gtheta,
                                                             k > 0 ? a->fd_all + (size_t)(k - 1) * n : nullptr, nkn);
                std::swap(lam, lamo);
                std::swap(nkc, nkn);
            }
        }
        if (!a->fd_all && T > 0 && !a->side) {
            TN_HIP(hipStreamCreateWithFlags(&a->side, hipStreamNonBlocking));
            for (hipEvent_t* e : { &a->ev_start, &a->ev_ready[0], &a->ev_ready[1], &a->ev_free[0], &a->ev_free[1] }) {
                TN_HIP(hipEventCreateWithFlags(e, hipEventDisableTiming));
            }
        }
        // segment sg's state k (k0 <= k <= k1): its checkpoint, else buffer sg % 2
        auto state = [&](int sg, int k) {
            const int k0 = sg * a->ckpt;
            return k == k0 ? a->ckpts + (size_t)sg * n : a->seg[sg & 1] + (size_t)(k - k0) * n;
        };
        // recompute segment sg on the side stream: its states and each one's F, D
        bool used[2] = { false, false };
        auto recompute = [&](int sg) {
            const int b = sg & 1, k0 = sg * a->ckpt, k1 = std::min(T, k0 + a->ckpt);
            if (used[b]) {   // the backward that read this buffer last is done with it
                TN_HIP(hipStreamWaitEvent(a->side, a->ev_free[b], 0));
            }
            used[b] = true;
            for (int k = k0; k < k1; ++k) {
                tnadj::k_adj_step<<<nb, tnk::kBlock, 0, a->side>>>(A, state(sg, k), state(sg, k + 1),
                                                                   a->fd[b] + (size_t)(k - k0) * n);
            }
            TN_HIP(hipEventRecord(a->ev_ready[b], a->side));
        };
        if (!a->fd_all && T > 0) {
            TN_HIP(hipEventRecord(a->ev_start, s));   // the side stream starts after the work queued before
            TN_HIP(hipStreamWaitEvent(a->side, a->ev_start, 0));
            recompute(a->nseg - 1);
        }
        for (int sg = (a->fd_all || T == 0) ? -1 : a->nseg - 1; sg >= 0; --sg) {
            const int b = sg & 1, k0 = sg * a->ckpt, k1 = std::min(T, k0 + a->ckpt);   // states k0 .. k1-1 here
            if (sg > 0) {   // the segment before, beside this one's backward
                recompute(sg - 1);
            }
            TN_HIP(hipStreamWaitEvent(s, a->ev_ready[b], 0));
            if (sg == a->nseg - 1) {   // X_T: the loss gradient starts the adjoint
                tnadj::k_adj_loss_grad<<<nb, tnk::kBlock, 0, s>>>(A, state(sg, k1), lam, a->gtheta);
            }
            float4 *nkc = a->nk, *nkn = a->nk2;
            // the segment's last step; the steps before get theirs from k_adj_back
            tnadj::k_adj_nu<<<nb, tnk::kBlock, 0, s>>>(A, a->fd[b] + (size_t)(k1 - 1 - k0) * n, lam, nkc);
            for (int k = k1 - 1; k >= k0; --k) {
                const int r = k - k0;
                tnadj::k_adj_back<<<nb, tnk::kBlock, 0, s>>>(A, state(sg, k), nkc, lam, lamo, a->gtheta,
                                                             k > k0 ? a->fd[b] + (size_t)(r - 1) * n : nullptr, nkn);
                std::swap(lam, lamo);
                std::swap(nkc, nkn);
            }
            TN_HIP(hipEventRecord(a->ev_free[b], s));
        }
        if (grad_x0) {
            tnadj::k_adj_unload<<<nb, tnk::kBlock, 0, s>>>(n, nullptr, lam, grad_x0);
        }
        if (grad_theta) {
            tnadj::k_adj_theta_out<<<blocks_for(a->gcells, tnk::kBlock), tnk::kBlock, 0, s>>>(a->gcells, a->gtheta,
                                                                                             grad_theta);
        }
        TN_HIP(hipGetLastError());
        TN_HIP(hipStreamSynchronize(s));
    });
}

int tn_adj_loss_terms(const tn_adj* a, double* terms) {
    return guarded([&] {
        arg_check(a != nullptr && terms != nullptr, "adjoint: null argument");
        for (int k = 0; k < 3; ++k) {
            terms[k] = a->loss[k];
        }
    });
}

int tn_adj_destroy(tn_adj* a) {
    return guarded([&] {
        if (a) {
            (void)hipSetDevice(a->device);
            delete a;
        }
    });
}

}  // extern "C"
