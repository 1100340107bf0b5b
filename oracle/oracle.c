/*
 * oracle.c -- CPU restatement of v2mesh's relaxation (see oracle.h).
 *
 * TEST INFRASTRUCTURE ONLY: the checker of the HIP path, never linked by it.
 *
 * Each function names the reference lines it restates. The arithmetic keeps
 * the reference's association and operand order exactly (the results must be
 * bit-identical), compiled with -ffp-contract=off like the reference
 * (CMakeLists.txt:69-72). Label volumes (gm == 0), gray-scale volumes
 * (gm = m > 0 thresholds) and shape fields (gm < 0, the labels' analytic-field
 * program) are covered, the voxel trap included.
 */
#include "oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

static _Thread_local char g_err[256];

const char* orc_last_error(void) {
    return g_err;
}

static int fail(const char* msg) {
    snprintf(g_err, sizeof g_err, "%s", msg);
    return 1;
}

/* ---- the field ------------------------------------------------------------------ */

typedef struct {
    int nx, ny, nz, nbx, nby;
    float vx, vy, vz;
    const uint16_t* L;
    const int32_t* bl_cnt;
    const uint16_t* bl_lab;
    const int32_t* bl_slot;
    const float* phi;
    const float* h;
    const float* gI;    /* gray-scale mode: intensity per voxel */
    const float* gTW;   /* [t_0..t_{m-1}, W_0..W_{m-1}] */
    int gm;             /* m thresholds (0: label volume) */
} Fld;

static Fld fld_of(const tn_grid_view* g) {
    Fld f;
    f.nx = g->nx;
    f.ny = g->ny;
    f.nz = g->nz;
    f.nbx = g->nbx;
    f.nby = g->nby;
    f.vx = g->vs[0];
    f.vy = g->vs[1];
    f.vz = g->vs[2];
    f.L = g->L;
    f.bl_cnt = g->bl_cnt;
    f.bl_lab = g->bl_lab;
    f.bl_slot = g->bl_slot;
    f.phi = g->phi;
    f.h = g->h;
    f.gI = g->gI;
    f.gTW = g->gTW;
    f.gm = g->gm;
    return f;
}

static inline size_t vox(const Fld* f, int i, int j, int k) {
    return (size_t)i + (size_t)f->nx * ((size_t)j + (size_t)f->ny * (size_t)k);
}

/* v2m_brick_of, src/opencl/v2m_grid_body.cl:116-118 */
static inline int brick_at(const Fld* f, int i, int j, int k) {
    return (i >> 3) + f->nbx * ((j >> 3) + f->nby * (k >> 3));
}

/* v2m_slot_of, src/opencl/v2m_grid_body.cl:122-134: the slot of label l in brick b
 * (-1 if the brick is uniform or lacks l); *has = l is one of the brick's labels */
static inline int slot_in(const Fld* f, int b, int l, int* has) {
    const int cnt = f->bl_cnt[b];
    *has = 0;
    for (int s = 0; s < cnt && s < TN_BL; ++s) {
        if (f->bl_lab[(size_t)b * TN_BL + s] == l) {
            *has = 1;
            return f->bl_slot[b] < 0 ? -1 : f->bl_slot[b] + s;
        }
    }
    return -1;
}

/* v2m_phi_vox, src/opencl/v2m_grid_body.cl:137-153 */
static float phi_voxel(const Fld* f, int l, int i, int j, int k) {
    if (i < 0 || j < 0 || k < 0 || i >= f->nx || j >= f->ny || k >= f->nz) {
        return l == 0 ? 1.0f : 0.0f;
    }
    int has;
    const int s = slot_in(f, brick_at(f, i, j, k), l, &has);
    if (s < 0) {
        return (has && f->L[vox(f, i, j, k)] == l) ? 1.0f : 0.0f;
    }
    return f->phi[(size_t)s * TN_SLOT + (size_t)((i & 7) + TN_BS * ((j & 7) + TN_BS * (k & 7)))];
}

/* v2m_phi_cell, src/opencl/v2m_grid_body.cl:159-186: the 8 corners of the cell at
 * (i0,j0,k0), corner bit 0 -> +i, bit 1 -> +j, bit 2 -> +k */
static void phi_corners(const Fld* f, int l, int i0, int j0, int k0, float* c) {
    if (i0 >= 0 && j0 >= 0 && k0 >= 0 && i0 + 1 < f->nx && j0 + 1 < f->ny && k0 + 1 < f->nz &&
        (i0 >> 3) == ((i0 + 1) >> 3) && (j0 >> 3) == ((j0 + 1) >> 3) && (k0 >> 3) == ((k0 + 1) >> 3)) {
        int has;
        const int s = slot_in(f, brick_at(f, i0, j0, k0), l, &has);
        if (s >= 0) {
            const float* base = f->phi + (size_t)s * TN_SLOT + (size_t)((i0 & 7) + TN_BS * ((j0 & 7) + TN_BS * (k0 & 7)));
            for (int m = 0; m < 8; ++m) {
                c[m] = base[(m & 1) + TN_BS * ((m >> 1) & 1) + TN_BS * TN_BS * (m >> 2)];
            }
        } else {
            for (int m = 0; m < 8; ++m) {
                c[m] = (has && f->L[vox(f, i0 + (m & 1), j0 + ((m >> 1) & 1), k0 + (m >> 2))] == l) ? 1.0f : 0.0f;
            }
        }
        return;
    }
    for (int m = 0; m < 8; ++m) {
        c[m] = phi_voxel(f, l, i0 + (m & 1), j0 + ((m >> 1) & 1), k0 + ((m >> 2) & 1));
    }
}

/* trilinear interpolation of 8 corner values (the shared tail of v2m_phi_at,
 * src/opencl/v2m_grid_body.cl:196-199, and v2m_psi, src/opencl/v2m_seed_body.cl:201-204) */
static inline float trilerp(const float* c, float fx, float fy, float fz) {
    const float x00 = c[0] + fx * (c[1] - c[0]), x10 = c[2] + fx * (c[3] - c[2]);
    const float x01 = c[4] + fx * (c[5] - c[4]), x11 = c[6] + fx * (c[7] - c[6]);
    const float y0 = x00 + fy * (x10 - x00), y1 = x01 + fy * (x11 - x01);
    return y0 + fz * (y1 - y0);
}

/* ---- shape fields (gm < 0): the labels' analytic-field program ------------------ */
/* The program layout and op codes are v2mesh's (src/opencl/v2m_sdf_body.cl:1-45):
 * [0] labels N, [1] field width w, [2..4] origin, [5..5+N-1] code offsets, [5+N]
 * crease blend radius, [6+N] gap-closing radius, [7+N] cull margin, [8+N] brick
 * table offset (0: none); code: an op, its operands; op 0 ends a label's code. */
enum { SDF_END = 0, SDF_PRIM = 1, SDF_MAX = 2, SDF_MIN = 3, SDF_NEG = 4, SDF_CONST = 5, SDF_BBOX = 6,
       SDF_MCONST = 7, SDF_ADD = 8, SDF_SCALE = 9, SDF_CMAX = 10, SDF_CMIN = 11 };
enum { SDF_SPHERE = 1, SDF_BOX = 2, SDF_CYL = 3, SDF_PLANE = 4, SDF_SLAB = 5, SDF_TORUS = 6, SDF_ELLIPSOID = 7,
       SDF_CONE = 8, SDF_INFCYL = 9 };
#define SDF_STACK 12

typedef struct {
    float v, x, y, z;   /* a value (> 0 inside) and its gradient */
} SdfV;

/* parameters per primitive type (v2m_sdf_nparam, :47-74) */
static int sdf_nparam(int type) {
    switch (type) {
    case SDF_SPHERE: return 4;
    case SDF_BOX: case SDF_PLANE: return 6;
    case SDF_CYL: case SDF_INFCYL: return 7;
    case SDF_SLAB: case SDF_TORUS: case SDF_CONE: return 8;
    case SDF_ELLIPSOID: return 15;
    default: return 0;
    }
}

/* one primitive's signed distance at (x, y, z), > 0 inside (v2m_sdf_prim, :76-161) */
static float sdf_prim(int type, const float* q, float x, float y, float z) {
    if (type == SDF_SPHERE) {
        const float dx = x - q[0], dy = y - q[1], dz = z - q[2];
        return q[3] - sqrtf(dx * dx + dy * dy + dz * dz);
    }
    if (type == SDF_BOX) {
        const float cx = 0.5f * (q[0] + q[3]), cy = 0.5f * (q[1] + q[4]), cz = 0.5f * (q[2] + q[5]);
        const float ux = fabsf(x - cx) - 0.5f * fabsf(q[3] - q[0]), uy = fabsf(y - cy) - 0.5f * fabsf(q[4] - q[1]),
                    uz = fabsf(z - cz) - 0.5f * fabsf(q[5] - q[2]);
        const float ox = fmaxf(ux, 0.0f), oy = fmaxf(uy, 0.0f), oz = fmaxf(uz, 0.0f);
        const float out = sqrtf(ox * ox + oy * oy + oz * oz), in = fminf(fmaxf(ux, fmaxf(uy, uz)), 0.0f);
        return -(out + in);
    }
    if (type == SDF_PLANE) {
        return -((x - q[0]) * q[3] + (y - q[1]) * q[4] + (z - q[2]) * q[5]);
    }
    if (type == SDF_SLAB) {
        const float t = (x - q[0]) * q[3] + (y - q[1]) * q[4] + (z - q[2]) * q[5];
        return fminf(t - q[6], q[7] - t);
    }
    if (type == SDF_INFCYL) {
        const float dx = x - q[0], dy = y - q[1], dz = z - q[2];
        const float t = dx * q[3] + dy * q[4] + dz * q[5];
        const float rx = dx - t * q[3], ry = dy - t * q[4], rz = dz - t * q[5];
        return q[6] - sqrtf(rx * rx + ry * ry + rz * rz);
    }
    if (type == SDF_TORUS) {
        const float dx = x - q[0], dy = y - q[1], dz = z - q[2];
        const float hh = dx * q[3] + dy * q[4] + dz * q[5];
        const float rx = dx - hh * q[3], ry = dy - hh * q[4], rz = dz - hh * q[5];
        const float a = sqrtf(rx * rx + ry * ry + rz * rz) - q[6];
        return q[7] - sqrtf(a * a + hh * hh);
    }
    if (type == SDF_CYL) {   /* capped cylinder, I. Quilez's exact form */
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
    if (type == SDF_CONE) {   /* capped frustum, I. Quilez's exact form */
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
    if (type == SDF_ELLIPSOID) {   /* the bound k0 (k0 - 1) / k1, exact on the surface */
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

/* min (sgn 1) or max (sgn -1) of a and b with the crease blend of radius k and the
 * gap closing of radius kgap (v2m_sdf_minmax, :191-273) */
static SdfV sdf_minmax(SdfV a, SdfV b, float sgn, float k, float kgap, int grad, int* near) {
    const float va = sgn * a.v, vb = sgn * b.v;
    const int sliver = kgap > 0.0f && va > -kgap && vb > -kgap && va + vb < kgap;
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
    SdfV r = a;
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
        r.v = sgn * (vb * (1.0f - h) + va * h - k * h * (1.0f - h));
        r.x = h * a.x + (1.0f - h) * b.x;
        r.y = h * a.y + (1.0f - h) * b.y;
        r.z = h * a.z + (1.0f - h) * b.z;
    } else if (vb < va) {
        r = b;
    }
    r.v -= sgn * pen;
    if (dpen > 0.0f) {
        r.x += dpen * (a.x + b.x);
        r.y += dpen * (a.y + b.y);
        r.z += dpen * (a.z + b.z);
    }
    return r;
}

static int sdf_is_minmax(int op) {
    return op == SDF_MAX || op == SDF_MIN || op == SDF_CMAX || op == SDF_CMIN;
}

/* label l's field at grid-mm point p and (wantg) its gradient: the postfix program
 * on a stack of SDF_STACK entries, values first and the gradients only where a gap
 * or a crease needs them (v2m_sdf_evalv, :275-478) */
static SdfV sdf_evalv(const float* prog, int l, float px, float py, float pz, int wantg) {
    const int N = (int)prog[0];
    const SdfV zero = { 0.0f, 0.0f, 0.0f, 0.0f };
    SdfV none = zero;
    none.v = -1e30f;
    if (l < 0 || l >= N) {
        return none;
    }
    const float x = px + prog[2], y = py + prog[3], z = pz + prog[4];
    int start = (int)prog[5 + l];
    const int T = (int)prog[8 + N];
    if (T > 0) {   /* the point's brick's own program */
        const int nbx = (int)prog[T], nby = (int)prog[T + 1], nbz = (int)prog[T + 2];
        const float vx = prog[T + 3];
        int bi = (int)floorf(px / vx + 0.5f) >> 3, bj = (int)floorf(py / vx + 0.5f) >> 3,
            bk = (int)floorf(pz / vx + 0.5f) >> 3;
        bi = bi < 0 ? 0 : (bi >= nbx ? nbx - 1 : bi);
        bj = bj < 0 ? 0 : (bj >= nby ? nby - 1 : bj);
        bk = bk < 0 ? 0 : (bk >= nbz ? nbz - 1 : bk);
        start = (int)prog[(int)prog[T + 4 + bi + nbx * (bj + nby * bk)] + l];
    }
    SdfV st[SDF_STACK];   /* st[sp - 1] is the top */
    int sp = 0;
    const float kgap = prog[6 + N], kblend = prog[5 + N];
    int grad = wantg != 0, near = 0;
    for (int pass = 0; pass < 2; ++pass) {
        sp = 0;
        int pc = start;
        near = 0;
        for (int guard = 0; guard < 100000; ++guard) {
            const int op = (int)prog[pc];
            if (op == SDF_END) {
                break;
            }
            if (sdf_is_minmax(op)) {   /* second op top -> top, pop */
                if (sp > 1) {
                    st[sp - 2] = sdf_minmax(st[sp - 2], st[sp - 1], (op == SDF_MAX || op == SDF_CMAX) ? -1.0f : 1.0f,
                                            kblend, op >= SDF_CMAX ? 0.0f : kgap, grad, &near);
                    --sp;
                }
                pc += 1;
                continue;
            }
            if (op == SDF_NEG) {
                if (sp > 0) {
                    st[sp - 1].v = -st[sp - 1].v;
                    st[sp - 1].x = -st[sp - 1].x;
                    st[sp - 1].y = -st[sp - 1].y;
                    st[sp - 1].z = -st[sp - 1].z;
                }
                pc += 1;
                continue;
            }
            if (op == SDF_ADD) {   /* top += second, pop */
                if (sp > 1) {
                    SdfV t = st[sp - 1];
                    t.v += st[sp - 2].v;
                    t.x += st[sp - 2].x;
                    t.y += st[sp - 2].y;
                    t.z += st[sp - 2].z;
                    st[sp - 2] = t;
                    --sp;
                }
                pc += 1;
                continue;
            }
            if (op == SDF_SCALE) {
                const float fs = prog[pc + 1];
                if (sp > 0) {
                    st[sp - 1].v *= fs;
                    st[sp - 1].x *= fs;
                    st[sp - 1].y *= fs;
                    st[sp - 1].z *= fs;
                }
                pc += 2;
                continue;
            }
            SdfV nv = zero;
            if (op == SDF_PRIM) {   /* the value and (grad) this primitive's central differences */
                const int type = (int)prog[pc + 1];
                const float* q = prog + pc + 2;
                const float e = fmaxf(1e-4f * (fabsf(q[0]) + fabsf(q[1]) + fabsf(q[2]) + 1.0f), 1e-5f);
                const int ne = grad ? 7 : 1;
                float fv[7];
                for (int k = 0; k < ne; ++k) {
                    const float dd = k == 0 ? 0.0f : ((k & 1) ? e : -e);
                    const int ax = (k - 1) >> 1;
                    fv[k] = sdf_prim(type, q, x + (ax == 0 ? dd : 0.0f), y + (ax == 1 ? dd : 0.0f),
                                     z + (ax == 2 ? dd : 0.0f));
                }
                nv.v = fv[0];
                if (grad) {
                    nv.x = (fv[1] - fv[2]) / (2.0f * e);
                    nv.y = (fv[3] - fv[4]) / (2.0f * e);
                    nv.z = (fv[5] - fv[6]) / (2.0f * e);
                }
                pc += 2 + sdf_nparam(type);
            } else if (op == SDF_BBOX) {   /* outside the box (+ margin): the bound, code skipped */
                const float m = prog[7 + N];
                const float* q = prog + pc + 1;
                if (!(x < q[0] - m || y < q[1] - m || z < q[2] - m || x > q[3] + m || y > q[4] + m || z > q[5] + m)) {
                    pc += 9;
                    continue;
                }
                nv.v = q[7] * m;
                pc += 9 + (int)q[6];
            } else if (op == SDF_MCONST) {
                nv.v = prog[pc + 1] * prog[7 + N];
                pc += 2;
            } else if (op == SDF_CONST) {
                nv.v = prog[pc + 1];
                pc += 2;
            } else {
                break;   /* a malformed program */
            }
            {   /* an operand followed by a min / max (after an optional NEG) combines at once */
                const int op1 = (int)prog[pc], neg = op1 == SDF_NEG;
                const int op2 = neg ? (int)prog[pc + 1] : op1;
                if (sp > 0 && sdf_is_minmax(op2)) {
                    if (neg) {
                        nv.v = -nv.v;
                        nv.x = -nv.x;
                        nv.y = -nv.y;
                        nv.z = -nv.z;
                    }
                    st[sp - 1] = sdf_minmax(st[sp - 1], nv, (op2 == SDF_MAX || op2 == SDF_CMAX) ? -1.0f : 1.0f, kblend,
                                            op2 >= SDF_CMAX ? 0.0f : kgap, grad, &near);
                    pc += 1 + neg;
                    continue;
                }
            }
            if (sp < SDF_STACK) {
                st[sp++] = nv;
            }
        }
        if (grad || !near) {
            break;
        }
        grad = 1;
    }
    return sp > 0 ? st[sp - 1] : none;
}

static float sdf_value(const float* prog, int l, float px, float py, float pz) {
    return sdf_evalv(prog, l, px, py, pz, 0).v;
}

/* v2m_phi_f, src/opencl/v2m_seed_body.cl:33-39: v2m_phi_at (src/opencl/v2m_grid_body.cl:
 * 189-200), or for shape fields clamp(0.5 + s_l / w) */
static float phi_point(const Fld* f, int l, float px, float py, float pz) {
    if (f->gm < 0) {
        return fminf(1.0f, fmaxf(0.0f, 0.5f + sdf_value(f->gTW, l, px, py, pz) / f->gTW[1]));
    }
    const float u = px / f->vx, v = py / f->vy, w = pz / f->vz;
    const int i0 = (int)floorf(u), j0 = (int)floorf(v), k0 = (int)floorf(w);
    float c[8];
    phi_corners(f, l, i0, j0, k0, c);
    return trilerp(c, u - (float)i0, v - (float)j0, w - (float)k0);
}

/* v2m_gray_at, src/opencl/v2m_seed_body.cl:65-93: the trilinear intensity at p and
 * (g non-null) its gradient; a corner outside the grid reads t_0 - W_0 */
static float gray_at(const Fld* f, float px, float py, float pz, float* g) {
    const float u = px / f->vx, v = py / f->vy, w = pz / f->vz;
    const int i0 = (int)floorf(u), j0 = (int)floorf(v), k0 = (int)floorf(w);
    const float fx = u - (float)i0, fy = v - (float)j0, fz = w - (float)k0;
    const float out = f->gTW[0] - f->gTW[f->gm];
    float c[8];
    for (int m = 0; m < 8; ++m) {
        const int i = i0 + (m & 1), j = j0 + ((m >> 1) & 1), k = k0 + (m >> 2);
        c[m] = (i < 0 || j < 0 || k < 0 || i >= f->nx || j >= f->ny || k >= f->nz) ? out : f->gI[vox(f, i, j, k)];
    }
    const float x00 = c[0] + fx * (c[1] - c[0]), x10 = c[2] + fx * (c[3] - c[2]);
    const float x01 = c[4] + fx * (c[5] - c[4]), x11 = c[6] + fx * (c[7] - c[6]);
    const float y0 = x00 + fy * (x10 - x00), y1 = x01 + fy * (x11 - x01);
    if (g) {
        const float dx0 = (c[1] - c[0]) + fy * ((c[3] - c[2]) - (c[1] - c[0]));
        const float dx1 = (c[5] - c[4]) + fy * ((c[7] - c[6]) - (c[5] - c[4]));
        g[0] = (dx0 + fz * (dx1 - dx0)) / f->vx;
        g[1] = ((x10 - x00) + fz * ((x11 - x01) - (x10 - x00))) / f->vy;
        g[2] = (y1 - y0) / f->vz;
    }
    return y0 + fz * (y1 - y0);
}

/* v2m_gray_psi, src/opencl/v2m_seed_body.cl:97-113: for adjacent labels (|a - b| = 1)
 * of a gray-scale volume, psi_ab = +-2 (I - t_k) / W_k with k = min(a, b), and its
 * gradient (g non-null). Returns 0 where it does not apply (also for a pair above
 * the last threshold, which a gray-scale volume does not hold: gTW has no entry). */
static int gray_psi(const Fld* f, int a, int b, float px, float py, float pz, float* v, float* g) {
    const int k = a < b ? a : b;
    if (f->gm <= 0 || (a - b != 1 && b - a != 1) || k >= f->gm) {
        return 0;
    }
    const float sg = (a > b ? 2.0f : -2.0f) / f->gTW[f->gm + k];
    *v = sg * (gray_at(f, px, py, pz, g) - f->gTW[k]);
    if (g) {
        g[0] *= sg;
        g[1] *= sg;
        g[2] *= sg;
    }
    return 1;
}

/* v2m_psi, src/opencl/v2m_seed_body.cl:179-205: phi_a - phi_b, trilinear (the
 * gray-scale form for adjacent labels) */
static float psi_point(const Fld* f, int a, int b, float px, float py, float pz) {
    float gv;
    if (f->gm < 0) {   /* shape fields: (s_a - s_b) / w */
        return (sdf_value(f->gTW, a, px, py, pz) - sdf_value(f->gTW, b, px, py, pz)) / f->gTW[1];
    }
    if (gray_psi(f, a, b, px, py, pz, &gv, NULL)) {
        return gv;
    }
    const float u = px / f->vx, v = py / f->vy, w = pz / f->vz;
    const int i0 = (int)floorf(u), j0 = (int)floorf(v), k0 = (int)floorf(w);
    float c[8], e[8];
    phi_corners(f, a, i0, j0, k0, c);
    phi_corners(f, b, i0, j0, k0, e);
    for (int m = 0; m < 8; ++m) {
        c[m] -= e[m];
    }
    return trilerp(c, u - (float)i0, v - (float)j0, w - (float)k0);
}

/* v2m_psi_grad, src/opencl/v2m_seed_body.cl:211-247: psi and its analytic gradient
 * from one trilinear cell (the gray-scale form for adjacent labels) */
static float psi_grad_point(const Fld* f, int a, int b, float px, float py, float pz, float* g) {
    float gv;
    if (f->gm < 0) {   /* shape fields: the analytic gradients */
        const float w = f->gTW[1];
        const SdfV sa = sdf_evalv(f->gTW, a, px, py, pz, 1), sb = sdf_evalv(f->gTW, b, px, py, pz, 1);
        g[0] = (sa.x - sb.x) / w;
        g[1] = (sa.y - sb.y) / w;
        g[2] = (sa.z - sb.z) / w;
        return (sa.v - sb.v) / w;
    }
    if (gray_psi(f, a, b, px, py, pz, &gv, g)) {
        return gv;
    }
    const float u = px / f->vx, v = py / f->vy, w = pz / f->vz;
    const int i0 = (int)floorf(u), j0 = (int)floorf(v), k0 = (int)floorf(w);
    const float fx = u - (float)i0, fy = v - (float)j0, fz = w - (float)k0;
    float c[8], e[8];
    phi_corners(f, a, i0, j0, k0, c);
    phi_corners(f, b, i0, j0, k0, e);
    for (int m = 0; m < 8; ++m) {
        c[m] -= e[m];
    }
    const float x00 = c[0] + fx * (c[1] - c[0]), x10 = c[2] + fx * (c[3] - c[2]);
    const float x01 = c[4] + fx * (c[5] - c[4]), x11 = c[6] + fx * (c[7] - c[6]);
    const float y0 = x00 + fy * (x10 - x00), y1 = x01 + fy * (x11 - x01);
    const float dx0 = (c[1] - c[0]) + fy * ((c[3] - c[2]) - (c[1] - c[0]));
    const float dx1 = (c[5] - c[4]) + fy * ((c[7] - c[6]) - (c[5] - c[4]));
    g[0] = (dx0 + fz * (dx1 - dx0)) / f->vx;
    g[1] = ((x10 - x00) + fz * ((x11 - x01) - (x10 - x00))) / f->vy;
    g[2] = (y1 - y0) / f->vz;
    return y0 + fz * (y1 - y0);
}

/* v2m_labels_near, src/opencl/v2m_seed_body.cl:42-59: the labels of p's brick */
static int brick_labels_at(const Fld* f, float px, float py, float pz, int* lab) {
    const int i = (int)floorf(px / f->vx + 0.5f), j = (int)floorf(py / f->vy + 0.5f), k = (int)floorf(pz / f->vz + 0.5f);
    if (i < 0 || j < 0 || k < 0 || i >= f->nx || j >= f->ny || k >= f->nz) {
        lab[0] = 0;
        return 1;
    }
    const int b = brick_at(f, i, j, k);
    const int n = f->bl_cnt[b] < TN_BL ? f->bl_cnt[b] : TN_BL;
    for (int s = 0; s < n; ++s) {
        lab[s] = f->bl_lab[(size_t)b * TN_BL + s];
    }
    return n;
}

/* v2m_label_of (voxmode 0), src/opencl/v2m_seed_body.cl:116-175: argmax of phi over
 * the brick's labels, plus the runner-up and the margin. A gray-scale volume labels p
 * by the number of thresholds <= I(p); the runner-up lies across the nearer threshold
 * in membership units (none above the last). */
static int owner_at(const Fld* f, float px, float py, float pz, int* second, float* margin) {
    if (f->gm > 0) {
        const float I = gray_at(f, px, py, pz, NULL);
        const int m = f->gm;
        int l = 0;
        while (l < m && I >= f->gTW[l]) {
            ++l;
        }
        const float dl = l > 0 ? (I - f->gTW[l - 1]) / f->gTW[m + l - 1] : 1e30f;
        const float du = l < m ? (f->gTW[l] - I) / f->gTW[m + l] : 1e30f;
        *second = dl < du ? l - 1 : (l < m ? l + 1 : TN_NOLAB);
        *margin = 2.0f * fminf(dl, du);
        return l;
    }
    int lab[TN_BL];
    const int n = brick_labels_at(f, px, py, pz, lab);
    int best = lab[0], sec = TN_NOLAB;
    float pb = f->gm < 0 ? -1e30f : -1.0f, ps = pb;
    if (n == 1) {
        *second = TN_NOLAB;
        *margin = 1.0f;
        return best;
    }
    for (int s = 0; s < n; ++s) {   /* (shape fields: ranked by s itself) */
        const float v = f->gm < 0 ? sdf_value(f->gTW, lab[s], px, py, pz) : phi_point(f, lab[s], px, py, pz);
        if (v > pb) {
            ps = pb;
            sec = best;
            pb = v;
            best = lab[s];
        } else if (v > ps) {
            ps = v;
            sec = lab[s];
        }
    }
    if (f->gm < 0) {
        pb = fminf(1.0f, fmaxf(0.0f, 0.5f + pb / f->gTW[1]));
        ps = sec == TN_NOLAB ? 0.0f : fminf(1.0f, fmaxf(0.0f, 0.5f + ps / f->gTW[1]));
    }
    *second = sec;
    *margin = pb - ps;
    return best;
}

/* v2m_bisect_dir, src/opencl/v2m_seed_body.cl:255-320: bracket psi_ab = 0 along
 * direction u (doubling probes from the Newton guess tg), then Illinois regula falsi */
static int bracket_root(const Fld* f, int a, int b, float* p, const float* u, float v0, float rad, float tg) {
    if (v0 == 0.0f) {
        return 1;
    }
    const float sg = v0 > 0.0f ? -1.0f : 1.0f;
    float t0 = 0.0f, t1 = -1.0f, f0 = v0, f1 = 0.0f;
    float t = (tg > 0.0f && tg < rad) ? tg : 0.25f * rad;
    for (int k = 0; k < 8; ++k) {
        const float v = psi_point(f, a, b, p[0] + sg * t * u[0], p[1] + sg * t * u[1], p[2] + sg * t * u[2]);
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
        return 0;
    }
    int side = 0;
    for (int it = 0; it < 8 && f1 != 0.0f && fabsf(t1 - t0) > 1e-4f * rad; ++it) {
        const float tm = (f0 != f1) ? t1 - f1 * (t1 - t0) / (f1 - f0) : 0.5f * (t0 + t1);
        const float v = psi_point(f, a, b, p[0] + sg * tm * u[0], p[1] + sg * tm * u[1], p[2] + sg * tm * u[2]);
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
    for (int k = 0; k < 3; ++k) {
        p[k] += sg * t1 * u[k];
    }
    return 1;
}

/* v2m_project1, src/opencl/v2m_seed_body.cl:324-335 */
static int onto_interface(const Fld* f, int a, int b, float* p, float rad) {
    float g[3];
    const float v = psi_grad_point(f, a, b, p[0], p[1], p[2], g);
    const float gn = sqrtf(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
    if (gn < 1e-9f) {
        return 0;
    }
    const float u[3] = { g[0] / gn, g[1] / gn, g[2] / gn };
    return bracket_root(f, a, b, p, u, v, rad, 1.5f * fabsf(v) / gn);
}

/* v2m_project2 (gm == 0), src/opencl/v2m_seed_body.cl:342-420: alternate projections
 * onto a|b along n_ab and onto a|c along n_ac's part orthogonal to n_ab */
static int onto_junction(const Fld* f, int a, int b, int c, float* p, float rad) {
    float q[3] = { p[0], p[1], p[2] };
    if (f->gm < 0) {
        /* shape fields: of the three pairs' constraints, solve the two with the most
         * independent gradients, their shared label first (src/opencl/v2m_seed_body.cl:345-375) */
        float gab[3], gac[3], gbc[3];
        psi_grad_point(f, a, b, q[0], q[1], q[2], gab);
        psi_grad_point(f, a, c, q[0], q[1], q[2], gac);
        psi_grad_point(f, b, c, q[0], q[1], q[2], gbc);
        const float xa[3] = { gab[1] * gac[2] - gab[2] * gac[1], gab[2] * gac[0] - gab[0] * gac[2],
                              gab[0] * gac[1] - gab[1] * gac[0] };
        const float xb[3] = { gab[1] * gbc[2] - gab[2] * gbc[1], gab[2] * gbc[0] - gab[0] * gbc[2],
                              gab[0] * gbc[1] - gab[1] * gbc[0] };
        const float xc[3] = { gac[1] * gbc[2] - gac[2] * gbc[1], gac[2] * gbc[0] - gac[0] * gbc[2],
                              gac[0] * gbc[1] - gac[1] * gbc[0] };
        const float na = xa[0] * xa[0] + xa[1] * xa[1] + xa[2] * xa[2], nb = xb[0] * xb[0] + xb[1] * xb[1] + xb[2] * xb[2],
                    nc = xc[0] * xc[0] + xc[1] * xc[1] + xc[2] * xc[2];
        if (nb > na && nb >= nc) {   /* pivot b */
            const int t = a;
            a = b;
            b = t;
        } else if (nc > na && nc > nb) {   /* pivot c */
            const int t = a;
            a = c;
            c = b;
            b = t;
        }
    }
    for (int it = 0; it < 3; ++it) {
        float g1[3], g2[3];
        if (!onto_interface(f, a, b, q, rad)) {
            return 0;
        }
        psi_grad_point(f, a, b, q[0], q[1], q[2], g1);
        const float v2 = psi_grad_point(f, a, c, q[0], q[1], q[2], g2);
        const float a11 = g1[0] * g1[0] + g1[1] * g1[1] + g1[2] * g1[2];
        if (a11 < 1e-18f) {
            return 0;
        }
        const float pr = (g1[0] * g2[0] + g1[1] * g2[1] + g1[2] * g2[2]) / a11;
        float u[3] = { g2[0] - pr * g1[0], g2[1] - pr * g1[1], g2[2] - pr * g1[2] };
        const float un = sqrtf(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
        if (un < 1e-9f) {
            return 0;
        }
        u[0] /= un;
        u[1] /= un;
        u[2] /= un;
        if (!bracket_root(f, a, c, q, u, v2, rad, 1.5f * fabsf(v2) / un)) {
            return 0;
        }
    }
    if (!onto_interface(f, a, b, q, rad)) {
        return 0;
    }
    const float m2 = (q[0] - p[0]) * (q[0] - p[0]) + (q[1] - p[1]) * (q[1] - p[1]) + (q[2] - p[2]) * (q[2] - p[2]);
    if (m2 > rad * rad) {
        return 0;
    }
    p[0] = q[0];
    p[1] = q[1];
    p[2] = q[2];
    return 1;
}

/* ---- per-node bodies --------------------------------------------------------------- */

typedef struct {
    float b0;
    int nlev;
    float ox, oy, oz;
    int dim[TN_MAXLEV][3];
    int off[TN_MAXLEV];
    int nkeys;
} Hash;

/* v2m_h_at, src/opencl/v2m_particle_body.cl:30-36: h at the nearest voxel, clamped */
static inline float size_at(const Fld* f, float px, float py, float pz) {
    int i = (int)floorf(px / f->vx + 0.5f), j = (int)floorf(py / f->vy + 0.5f), k = (int)floorf(pz / f->vz + 0.5f);
    i = i < 0 ? 0 : (i >= f->nx ? f->nx - 1 : i);
    j = j < 0 ? 0 : (j >= f->ny ? f->ny - 1 : j);
    k = k < 0 ? 0 : (k >= f->nz ? f->nz - 1 : k);
    return f->h[vox(f, i, j, k)];
}

/* v2m_level_of, src/opencl/v2m_particle_body.cl:38-46 */
static inline int level_for(const Hash* H, float R) {
    int L = 0;
    while (L + 1 < H->nlev && H->b0 * (float)(1 << L) < R) {
        ++L;
    }
    return L;
}

/* v2m_bin_coord, src/opencl/v2m_particle_body.cl:48-52 */
static inline int cell_coord(const Hash* H, int L, float x, float o, int ax) {
    const int c = (int)floorf((x - o) / (H->b0 * (float)(1 << L)));
    const int n = H->dim[L][ax];
    return c < 0 ? 0 : (c >= n ? n - 1 : c);
}

/* v2m_bin_key, src/opencl/v2m_particle_body.cl:54-58 */
static inline int cell_key(const Hash* H, int L, float px, float py, float pz) {
    const int cx = cell_coord(H, L, px, H->ox, 0), cy = cell_coord(H, L, py, H->oy, 1), cz = cell_coord(H, L, pz, H->oz, 2);
    return H->off[L] + cx + H->dim[L][0] * (cy + H->dim[L][1] * cz);
}

/* the hash set-up of relax_cpu, src/v2m_particles.cpp:738-761 */
static void make_hash(const tn_grid_view* g, const tn_relax_params* prm, Hash* H) {
    memset(H, 0, sizeof *H);
    H->b0 = (prm->t + prm->skin) * g->hmin;
    H->ox = -0.5f * g->vs[0];
    H->oy = -0.5f * g->vs[1];
    H->oz = -0.5f * g->vs[2];
    const float ext[3] = { (float)g->nx * g->vs[0], (float)g->ny * g->vs[1], (float)g->nz * g->vs[2] };
    H->nlev = 1;
    while (H->nlev < TN_MAXLEV && H->b0 * (float)(1 << (H->nlev - 1)) < (prm->t + prm->skin) * g->hmax) {
        ++H->nlev;
    }
    int nkeys = 0;
    for (int L = 0; L < H->nlev; ++L) {
        H->off[L] = nkeys;
        const float bL = H->b0 * (float)(1 << L);
        for (int a = 0; a < 3; ++a) {
            const int d = (int)ceilf(ext[a] / bL);
            H->dim[L][a] = d > 1 ? d : 1;
        }
        nkeys += H->dim[L][0] * H->dim[L][1] * H->dim[L][2];
    }
    H->nkeys = nkeys;
}

void orc_hash_geometry(const tn_grid_view* g, const tn_relax_params* prm, tn_hash_info* out) {
    Hash H;
    make_hash(g, prm, &H);
    memset(out, 0, sizeof *out);
    out->b0 = H.b0;
    out->ox = H.ox;
    out->oy = H.oy;
    out->oz = H.oz;
    out->nlev = H.nlev;
    memcpy(out->dim, H.dim, sizeof H.dim);
    memcpy(out->off, H.off, sizeof H.off);
    out->nkeys = H.nkeys;
}

typedef struct {
    const float* hn;
    const float* P;
    const uint16_t* lab;
    const uint8_t* typ;
    const int32_t* cstart;
    const int32_t* sorted;
    const float* Ps;
    float t, skin;
} NbrIn;

/* v2m_neighbors, src/opencl/v2m_particle_body.cl:73-151: the K nearest valid
 * neighbours of node i in (distance, id) order, scanning levels Li-1..Li+1 */
static void truss_of(const Hash* H, const NbrIn* in, int i, int32_t* nbr, int32_t* nnb) {
    float ds[TN_K];
    int ids[TN_K];
    const float px = in->P[3 * i], py = in->P[3 * i + 1], pz = in->P[3 * i + 2];
    const float hi = in->hn[i];
    const float Ri = (in->t + in->skin) * hi;
    const int ti_int = in->typ[i] == TN_INTERIOR;
    const int li = in->lab[i];
    const int Li = level_for(H, Ri);
    int n = 0;
    for (int L = Li - 1; L <= Li + 1; ++L) {
        if (L < 0 || L >= H->nlev) {
            continue;
        }
        const float bL = H->b0 * (float)(1 << L);
        const float r = 0.5f * (Ri + bL);
        const int x0 = cell_coord(H, L, px - r, H->ox, 0), x1 = cell_coord(H, L, px + r, H->ox, 0);
        const int y0 = cell_coord(H, L, py - r, H->oy, 1), y1 = cell_coord(H, L, py + r, H->oy, 1);
        const int z0 = cell_coord(H, L, pz - r, H->oz, 2), z1 = cell_coord(H, L, pz + r, H->oz, 2);
        for (int cz = z0; cz <= z1; ++cz) {
            for (int cy = y0; cy <= y1; ++cy) {
                for (int cx = x0; cx <= x1; ++cx) {
                    const int key = H->off[L] + cx + H->dim[L][0] * (cy + H->dim[L][1] * cz);
                    for (int s = in->cstart[key]; s < in->cstart[key + 1]; ++s) {
                        const int j = in->sorted[s];
                        if (j == i) {
                            continue;
                        }
                        const float ex = in->Ps[4 * (size_t)s] - px, ey = in->Ps[4 * (size_t)s + 1] - py,
                                    ez = in->Ps[4 * (size_t)s + 2] - pz;
                        const float d2 = ex * ex + ey * ey + ez * ez;
                        const float hj = in->Ps[4 * (size_t)s + 3];
                        const float rj = (in->t + in->skin) * 0.5f * (hi + hj);
                        if (d2 >= rj * rj) {
                            continue;
                        }
                        if (ti_int && in->typ[j] == TN_INTERIOR && in->lab[j] != li) {
                            continue;
                        }
                        const float dist = sqrtf(d2);
                        if (n == TN_K && (dist > ds[TN_K - 1] || (dist == ds[TN_K - 1] && j > ids[TN_K - 1]))) {
                            continue;
                        }
                        int k = n < TN_K ? n++ : TN_K - 1;
                        while (k > 0 && (ds[k - 1] > dist || (ds[k - 1] == dist && ids[k - 1] > j))) {
                            ds[k] = ds[k - 1];
                            ids[k] = ids[k - 1];
                            --k;
                        }
                        ds[k] = dist;
                        ids[k] = j;
                    }
                }
            }
        }
    }
    nnb[i] = n;
    for (int k = 0; k < n; ++k) {
        nbr[(size_t)i * TN_K + k] = ids[k];
    }
}

/* v2m_force, src/opencl/v2m_particle_body.cl:161-184: repulsive bars in list order */
static void bar_force(const float* hn, const float* P, const uint8_t* typ, const int32_t* nbr, const int32_t* nnb,
                      float fscale, float fsurf, int i, float* F) {
    const float px = P[3 * i], py = P[3 * i + 1], pz = P[3 * i + 2];
    const float hi = hn[i];
    F[0] = F[1] = F[2] = F[3] = 0.0f;
    for (int k = 0; k < nnb[i]; ++k) {
        const int j = nbr[(size_t)i * TN_K + k];
        const float ex = px - P[3 * j], ey = py - P[3 * j + 1], ez = pz - P[3 * j + 2];
        const float l = sqrtf(ex * ex + ey * ey + ez * ez);
        const float hj = hn[j];
        const float l0 = ((typ[i] != TN_INTERIOR && typ[j] != TN_INTERIOR) ? fsurf : fscale) * 0.5f * (hi + hj);
        if (l >= l0 || l <= 0.0f) {
            continue;
        }
        const float fr = (l0 - l) / l;
        F[0] += fr * ex;
        F[1] += fr * ey;
        F[2] += fr * ez;
        F[3] += 1.0f;
    }
}

/* v2m_crossing, src/opencl/v2m_particle_body.cl:188-212 */
static void segment_crossing(const Fld* f, int a, int b, const float* p, const float* q, float* c) {
    float t0 = 0.0f, t1 = 1.0f;
    float v0 = psi_point(f, a, b, p[0], p[1], p[2]);
    float v1 = psi_point(f, a, b, q[0], q[1], q[2]);
    for (int it = 0; it < 6; ++it) {
        const float tm = 0.5f * (t0 + t1);
        const float vm = psi_point(f, a, b, p[0] + tm * (q[0] - p[0]), p[1] + tm * (q[1] - p[1]), p[2] + tm * (q[2] - p[2]));
        if (vm > 0.0f) {
            t0 = tm;
            v0 = vm;
        } else {
            t1 = tm;
            v1 = vm;
        }
    }
    const float t = (v0 - v1) != 0.0f ? t0 + (t1 - t0) * v0 / (v0 - v1) : 0.5f * (t0 + t1);
    for (int k = 0; k < 3; ++k) {
        c[k] = p[k] + t * (q[k] - p[k]);
    }
}

/* v2m_tangential, src/opencl/v2m_particle_body.cl:215-229 */
static void drop_normal(const Fld* f, int a, int b, const float* p, float* v) {
    float g[3];
    psi_grad_point(f, a, b, p[0], p[1], p[2], g);
    const float gn2 = g[0] * g[0] + g[1] * g[1] + g[2] * g[2];
    if (gn2 < 1e-12f) {
        return;
    }
    const float vn = (v[0] * g[0] + v[1] * g[1] + v[2] * g[2]) / gn2;
    for (int k = 0; k < 3; ++k) {
        v[k] -= vn * g[k];
    }
}

/* v2m_third_label, src/opencl/v2m_particle_body.cl:232-260 */
static int competing_label(const Fld* f, int a, int b, const float* q) {
    int nl[TN_BL];
    const int n = brick_labels_at(f, q[0], q[1], q[2], nl);
    int other = 0;
    for (int s = 0; s < n; ++s) {
        other |= nl[s] != a && nl[s] != b;
    }
    if (!other) {
        return TN_NOLAB;
    }
    const float pa = phi_point(f, a, q[0], q[1], q[2]);
    int c = TN_NOLAB;
    float pc = 0.5f * pa;
    for (int s = 0; s < n; ++s) {
        if (nl[s] != a && nl[s] != b) {
            const float w = phi_point(f, nl[s], q[0], q[1], q[2]);
            if (w > pc) {
                pc = w;
                c = nl[s];
            }
        }
    }
    return c;
}

/* v2m_valid_on, src/opencl/v2m_particle_body.cl:266-291 */
static int labels_dominate(const Fld* f, int a, int b, int c, const float* q) {
    int nl[TN_BL];
    const int n = brick_labels_at(f, q[0], q[1], q[2], nl);
    int other = 0;
    for (int s = 0; s < n; ++s) {
        other |= nl[s] != a && nl[s] != b && nl[s] != c;
    }
    if (!other) {
        return 1;
    }
    float mine = fminf(phi_point(f, a, q[0], q[1], q[2]), phi_point(f, b, q[0], q[1], q[2]));
    if (c != TN_NOLAB) {
        mine = fminf(mine, phi_point(f, c, q[0], q[1], q[2]));
    }
    for (int s = 0; s < n; ++s) {
        if (nl[s] != a && nl[s] != b && nl[s] != c && phi_point(f, nl[s], q[0], q[1], q[2]) > mine + 0.02f) {
            return 0;
        }
    }
    return 1;
}

typedef struct {
    float* P;
    const uint16_t* lab;
    uint8_t* typ;
    uint16_t* part;
    float* hn;
} NodeIO;

static inline float dist2(const float* q, const float* p) {
    return (q[0] - p[0]) * (q[0] - p[0]) + (q[1] - p[1]) * (q[1] - p[1]) + (q[2] - p[2]) * (q[2] - p[2]);
}

/* v2m_move_step (smooth mode), src/opencl/v2m_particle_body.cl:659-833: apply step s to
 * node i with interface trapping; returns |dp| / h (0 = stayed) */
static float trapped_step(const Fld* f, float h, float* s, float maxstep, float snap, int i, NodeIO* io) {
    float p[3], q[3];
    p[0] = io->P[3 * i];
    p[1] = io->P[3 * i + 1];
    p[2] = io->P[3 * i + 2];
    const int a = io->lab[i];
    int ty = io->typ[i];
    int b = io->part[2 * i], c = io->part[2 * i + 1];
    uint16_t* part = io->part;
    uint8_t* typ = io->typ;

    if (ty == TN_INTERIOR) {
        for (int k = 0; k < 3; ++k) {
            q[k] = p[k] + s[k];
        }
        int sec;
        float mg;
        const int l = owner_at(f, q[0], q[1], q[2], &sec, &mg);
        if (l != a) {
            b = l;
            float x[3], r[3];
            segment_crossing(f, a, b, p, q, x);
            for (int k = 0; k < 3; ++k) {
                r[k] = q[k] - x[k];
            }
            drop_normal(f, a, b, x, r);
            for (int k = 0; k < 3; ++k) {
                q[k] = x[k] + r[k];
            }
            if (!onto_interface(f, a, b, q, 0.5f * h)) {
                q[0] = x[0];
                q[1] = x[1];
                q[2] = x[2];
            }
            ty = TN_INTERFACE;
            part[2 * i] = (uint16_t)b;
        } else if (sec != TN_NOLAB && snap > 0.0f) {
            float gq[3];
            const float v = psi_grad_point(f, a, sec, q[0], q[1], q[2], gq);
            const float gn = sqrtf(gq[0] * gq[0] + gq[1] * gq[1] + gq[2] * gq[2]);
            if (gn > 1e-9f && v / gn < snap * h && onto_interface(f, a, sec, q, 1.2f * snap * h)) {
                ty = TN_INTERFACE;
                b = sec;
                part[2 * i] = (uint16_t)sec;
            }
        }
    } else if (ty == TN_INTERFACE) {
        drop_normal(f, a, b, p, s);
        for (int k = 0; k < 3; ++k) {
            q[k] = p[k] + s[k];
        }
        if (!onto_interface(f, a, b, q, 0.5f * h)) {
            return 0.0f;
        }
    } else {
        float g1[3], g2[3], tdir[3];
        psi_grad_point(f, a, b, p[0], p[1], p[2], g1);
        psi_grad_point(f, a, c, p[0], p[1], p[2], g2);
        tdir[0] = g1[1] * g2[2] - g1[2] * g2[1];
        tdir[1] = g1[2] * g2[0] - g1[0] * g2[2];
        tdir[2] = g1[0] * g2[1] - g1[1] * g2[0];
        const float tl2 = tdir[0] * tdir[0] + tdir[1] * tdir[1] + tdir[2] * tdir[2];
        const float st = tl2 > 1e-20f ? (s[0] * tdir[0] + s[1] * tdir[1] + s[2] * tdir[2]) / tl2 : 0.0f;
        for (int k = 0; k < 3; ++k) {
            q[k] = p[k] + st * tdir[k];
        }
        if (!onto_junction(f, a, b, c, q, 0.5f * h)) {
            return 0.0f;
        }
    }

    if (ty == TN_INTERFACE) {
        const int cc = competing_label(f, a, b, q);
        if (cc != TN_NOLAB) {
            float r[3] = { q[0], q[1], q[2] };
            if (onto_junction(f, a, b, cc, r, 0.5f * h)) {
                q[0] = r[0];
                q[1] = r[1];
                q[2] = r[2];
                ty = TN_JUNCTION;
                part[2 * i + 1] = (uint16_t)cc;
            }
        }
    }

    {
        const float lim = (maxstep + fmaxf(snap, 0.5f)) * h;
        const float m2 = dist2(q, p);
        if (!(m2 <= lim * lim)) {
            if (typ[i] == TN_INTERIOR) {
                part[2 * i] = TN_NOLAB;
            }
            return 0.0f;
        }
    }

    if (ty != TN_INTERIOR &&
        !labels_dominate(f, a, part[2 * i], ty >= TN_JUNCTION ? part[2 * i + 1] : TN_NOLAB, q)) {
        if (ty == TN_INTERFACE || typ[i] == TN_INTERIOR) {
            const int cc = competing_label(f, a, part[2 * i], q);
            float r[3] = { q[0], q[1], q[2] };
            if (cc != TN_NOLAB && onto_junction(f, a, part[2 * i], cc, r, 0.5f * h) &&
                labels_dominate(f, a, part[2 * i], cc, r)) {
                q[0] = r[0];
                q[1] = r[1];
                q[2] = r[2];
                ty = TN_JUNCTION;
                part[2 * i + 1] = (uint16_t)cc;
            } else {
                typ[i] = (uint8_t)(typ[i] == TN_INTERIOR ? TN_INTERIOR : ty);
                if (typ[i] == TN_INTERIOR) {
                    part[2 * i] = TN_NOLAB;
                }
                return 0.0f;
            }
        } else {
            return 0.0f;
        }
    }

    const float mv = sqrtf(dist2(q, p));
    io->P[3 * i] = q[0];
    io->P[3 * i + 1] = q[1];
    io->P[3 * i + 2] = q[2];
    typ[i] = (uint8_t)ty;
    io->hn[i] = size_at(f, q[0], q[1], q[2]);
    return mv / h;
}

/* ---- the voxel trap: the exact staircase of the label volume ----------------------- */

/* v2m_label_at, src/opencl/v2m_grid_body.cl:72-78 (0 outside the grid) */
static inline int label_at(const Fld* f, int i, int j, int k) {
    if (i < 0 || j < 0 || k < 0 || i >= f->nx || j >= f->ny || k >= f->nz) {
        return 0;
    }
    return f->L[vox(f, i, j, k)];
}

/* v2m_dda_hit, src/opencl/v2m_particle_body.cl:373-431: walk p -> p + s through the
 * voxels of label a; at the first face into another label, the face point into x,
 * that label into *b, and the ray fraction is returned; -1 if the step stays in a */
static float dda_hit(const Fld* f, int a, const float* p, const float* s, float* x, int* b) {
    const float vs[3] = { f->vx, f->vy, f->vz };
    float u[3], w[3];
    int id[3];
    for (int k = 0; k < 3; ++k) {
        u[k] = p[k] / vs[k] + 0.5f;
        w[k] = s[k] / vs[k];
        id[k] = (int)floorf(u[k]);
    }
    int l = label_at(f, id[0], id[1], id[2]);
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
        l = label_at(f, id[0], id[1], id[2]);
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

/* v2m_vox_nearest, src/opencl/v2m_particle_body.cl:437-529: the nearest point to q on
 * the a|b voxel faces (c == NOLAB) or the a|b|c voxel edges within R voxels; its
 * squared distance, or -1 */
static float vox_nearest(const Fld* f, int a, int b, int c, const float* q, int R, float* out) {
    const float vs[3] = { f->vx, f->vy, f->vz };
    float u[3];
    int ci[3];
    for (int k = 0; k < 3; ++k) {
        u[k] = q[k] / vs[k] + 0.5f;
        ci[k] = (int)floorf(u[k]);
    }
    float best = -1.0f;
    for (int vk = ci[2] - R; vk <= ci[2] + R; ++vk) {
        for (int vj = ci[1] - R; vj <= ci[1] + R; ++vj) {
            for (int vi = ci[0] - R; vi <= ci[0] + R; ++vi) {
                if (label_at(f, vi, vj, vk) != a) {
                    continue;
                }
                const int v[3] = { vi, vj, vk };
                for (int ax = 0; ax < 3; ++ax) {
                    const int o1 = (ax + 1) % 3, o2 = (ax + 2) % 3;
                    if (c == TN_NOLAB) {
                        for (int sd = 0; sd < 2; ++sd) {
                            int nn[3] = { v[0], v[1], v[2] };
                            nn[ax] += sd ? 1 : -1;
                            if (label_at(f, nn[0], nn[1], nn[2]) != b) {
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
                    } else {
                        for (int e = 0; e < 4; ++e) {
                            const int e1 = e & 1, e2 = e >> 1;
                            int hb = 0, hc = 0;
                            for (int m = 0; m < 4; ++m) {
                                int nn[3] = { v[0], v[1], v[2] };
                                nn[o1] += e1 - 1 + (m & 1);
                                nn[o2] += e2 - 1 + (m >> 1);
                                const int ln = label_at(f, nn[0], nn[1], nn[2]);
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

/* v2m_vox_third, src/opencl/v2m_particle_body.cl:533-545 */
static int vox_third(const Fld* f, int a, int b, const float* p) {
    const int i0 = (int)floorf(p[0] / f->vx), j0 = (int)floorf(p[1] / f->vy), k0 = (int)floorf(p[2] / f->vz);
    for (int m = 0; m < 8; ++m) {
        const int l = label_at(f, i0 + (m & 1), j0 + ((m >> 1) & 1), k0 + (m >> 2));
        if (l != a && l != b) {
            return l;
        }
    }
    return TN_NOLAB;
}

/* v2m_move_voxel, src/opencl/v2m_particle_body.cl:549-650: node i's step s (capped) in
 * voxel mode; returns |dp| / h */
static float voxel_step(const Fld* f, float h, float snap, int i, const float* s, NodeIO* io) {
    const float vmin = fminf(fminf(f->vx, f->vy), f->vz);
    float p[3], q[3], x[3];
    p[0] = io->P[3 * i];
    p[1] = io->P[3 * i + 1];
    p[2] = io->P[3 * i + 2];
    const int a = io->lab[i];
    int ty = io->typ[i], b = io->part[2 * i], c = io->part[2 * i + 1];
    const float sl = sqrtf(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
    const int R = 1 + (int)ceilf(sl / vmin);
    for (int k = 0; k < 3; ++k) {
        q[k] = p[k] + s[k];
    }
    if (ty == TN_INTERIOR) {
        int l;
        const float t = dda_hit(f, a, p, s, x, &l);
        if (t >= 0.0f) {
            b = l;
            if (vox_nearest(f, a, b, TN_NOLAB, q, R, q) < 0.0f) {
                q[0] = x[0];
                q[1] = x[1];
                q[2] = x[2];
            }
            ty = TN_INTERFACE;
        } else if (snap > 0.0f) {
            const int Rs = 1 + (int)ceilf(snap * h / vmin);
            float best = -1.0f, y[3];
            for (int dz = -1; dz <= 1; ++dz) {
                for (int dy = -1; dy <= 1; ++dy) {
                    for (int dx = -1; dx <= 1; ++dx) {
                        const int l2 = label_at(f, (int)floorf(q[0] / f->vx + 0.5f) + dx * Rs,
                                                (int)floorf(q[1] / f->vy + 0.5f) + dy * Rs,
                                                (int)floorf(q[2] / f->vz + 0.5f) + dz * Rs);
                        if (l2 == a || l2 == b) {
                            continue;
                        }
                        const float d2 = vox_nearest(f, a, l2, TN_NOLAB, q, Rs, y);
                        if (d2 >= 0.0f && d2 < snap * snap * h * h && (best < 0.0f || d2 < best)) {
                            best = d2;
                            b = l2;
                            x[0] = y[0];
                            x[1] = y[1];
                            x[2] = y[2];
                        }
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
    } else if (ty == TN_INTERFACE) {
        if (vox_nearest(f, a, b, TN_NOLAB, q, R, q) < 0.0f) {
            return 0.0f;
        }
    } else {
        if (vox_nearest(f, a, b, c, q, R, q) < 0.0f) {
            return 0.0f;
        }
    }
    if (ty == TN_INTERFACE) {
        const int cc = vox_third(f, a, b, q);
        if (cc != TN_NOLAB) {
            float r[3];
            const float d2 = vox_nearest(f, a, b, cc, q, 1, r);
            if (d2 >= 0.0f && d2 < 0.0625f * h * h) {
                q[0] = r[0];
                q[1] = r[1];
                q[2] = r[2];
                ty = TN_JUNCTION;
                c = cc;
            }
        }
    }
    const float mv = sqrtf(dist2(q, p));
    io->P[3 * i] = q[0];
    io->P[3 * i + 1] = q[1];
    io->P[3 * i + 2] = q[2];
    io->typ[i] = (uint8_t)ty;
    io->part[2 * i] = (uint16_t)(ty == TN_INTERIOR ? TN_NOLAB : b);
    io->part[2 * i + 1] = (uint16_t)(ty >= TN_JUNCTION ? c : TN_NOLAB);
    io->hn[i] = size_at(f, q[0], q[1], q[2]);
    return mv / h;
}

/* v2m_move, src/opencl/v2m_particle_body.cl:844-935: the step from the force (FIRE or
 * Jacobi), the settle and cap tests, the trapped step, and FIRE's velocity */
static float node_move(const Fld* f, const float* F, float dt, float maxstep, float snap, int voxmode, int i,
                       NodeIO* io, int fire, float* V, float fdt, float falpha, float* pw) {
    float s[3];
    const float p0 = io->P[3 * i], p1 = io->P[3 * i + 1], p2 = io->P[3 * i + 2];
    const float h = size_at(f, p0, p1, p2);
    const int ty0 = io->typ[i];
    if (fire) {
        pw[i] = 0.0f;
    }
    if (ty0 == TN_CORNER) {
        if (fire) {
            V[3 * i] = V[3 * i + 1] = V[3 * i + 2] = 0.0f;
        }
        return 0.0f;
    }
    const float stiff = F[4 * i + 3] > 1.0f ? F[4 * i + 3] : 1.0f;
    float ac[3];
    for (int k = 0; k < 3; ++k) {
        ac[k] = F[4 * i + k] / stiff;
    }
    if (fire) {
        float v[3] = { V[3 * i], V[3 * i + 1], V[3 * i + 2] };
        const float pv = ac[0] * v[0] + ac[1] * v[1] + ac[2] * v[2];
        pw[i] = pv;
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
            s[k] = dt * ac[k];
        }
    }
    const float sl = sqrtf(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
    if (sl < 5e-4f * h) {
        if (fire) {
            V[3 * i] = V[3 * i + 1] = V[3 * i + 2] = 0.0f;
        }
        return 0.0f;
    }
    if (sl > maxstep * h) {
        for (int k = 0; k < 3; ++k) {
            s[k] *= maxstep * h / sl;
        }
    }
    const float m = voxmode ? voxel_step(f, h, snap, i, s, io) : trapped_step(f, h, s, maxstep, snap, i, io);
    if (fire) {
        if (m > 0.0f && io->typ[i] == ty0) {
            V[3 * i] = (io->P[3 * i] - p0) / fdt;
            V[3 * i + 1] = (io->P[3 * i + 1] - p1) / fdt;
            V[3 * i + 2] = (io->P[3 * i + 2] - p2) / fdt;
        } else {
            V[3 * i] = V[3 * i + 1] = V[3 * i + 2] = 0.0f;
        }
    }
    return m;
}

/* ---- FIRE's global part: FireCtl, src/v2m_particles.h:47-76 ------------------------- */

typedef struct {
    float dt0, dt, dtmax, alpha;
    int npos, resets;
} Fire;

static void fire_init(Fire* c, const tn_relax_params* p) {
    c->dt0 = sqrtf(p->dt);
    c->dt = c->dt0;
    c->dtmax = p->fire_dtmax * c->dt0;
    c->alpha = 0.1f;
    c->npos = 0;
    c->resets = 0;
}

static int fire_may_stop(const Fire* c) {
    return c->dt >= 0.25f * c->dt0;
}

/* returns 1 when every velocity must be zeroed */
static int fire_update(Fire* c, double P) {
    if (P >= 0.0) {
        if (++c->npos > 5) {
            const float up = c->dt * 1.1f;
            c->dt = up < c->dtmax ? up : c->dtmax;
            c->alpha *= 0.99f;
        }
        return 0;
    }
    c->npos = 0;
    const float down = c->dt * 0.5f, lo = 0.05f * c->dt0;
    c->dt = down < lo ? lo : down;
    c->alpha = 0.1f;
    ++c->resets;
    return 1;
}

/* the histogram bin of src/v2m_particles.cpp:881 (std::log2 on a float is log2f) */
int32_t orc_hist_bin(float m) {
    if (!(m > 0.0f)) {
        return 0;
    }
    const int b = 20 + (int)floorf(log2f(m));
    return b < 0 ? 0 : (b > 31 ? 31 : b);
}

/* ---- the driver: relax_cpu, src/v2m_particles.cpp:728-976 ---------------------------- */

static int supported(const tn_grid_view* g, const tn_relax_params* prm) {
    if (!g || !prm) {
        return fail("oracle: null argument");
    }
    if (g->gm < 0 && (!g->gTW || g->ngTW < 9)) {
        return fail("oracle: a shape grid (gm < 0) needs its program in gTW");
    }
    if (g->gm > 0 && (!g->gI || !g->gTW || g->ngTW < 2 * (int64_t)g->gm)) {
        return fail("oracle: a gray-scale grid needs gI and 2 gm values in gTW");
    }
    return 0;
}

typedef struct {
    Hash H;
    int n;
    int32_t *key, *cstart, *sorted, *nbr, *nnb;
    float* Ps;
} Truss;

static void truss_free(Truss* T) {
    free(T->key);
    free(T->cstart);
    free(T->sorted);
    free(T->nbr);
    free(T->nnb);
    free(T->Ps);
}

static int truss_alloc(Truss* T, int n) {
    T->n = n;
    T->key = malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    T->cstart = malloc(sizeof(int32_t) * ((size_t)T->H.nkeys + 1));
    T->sorted = malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    T->nbr = calloc((size_t)(n > 0 ? n : 1) * TN_K, sizeof(int32_t));
    T->nnb = malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    T->Ps = malloc(sizeof(float) * 4 * (size_t)(n > 0 ? n : 1));
    if (!T->key || !T->cstart || !T->sorted || !T->nbr || !T->nnb || !T->Ps) {
        truss_free(T);
        return fail("oracle: out of memory");
    }
    return 0;
}

/* the rebuild lambda, src/v2m_particles.cpp:771-820 (without P0 = P) */
static void truss_rebuild(Truss* T, const Fld* f, const tn_relax_params* prm, const float* P, const uint16_t* lab,
                          const uint8_t* typ, const float* hn) {
    const int n = T->n;
    const Hash* H = &T->H;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
        const float h = size_at(f, P[3 * i], P[3 * i + 1], P[3 * i + 2]);
        T->key[i] = cell_key(H, level_for(H, (prm->t + prm->skin) * h), P[3 * i], P[3 * i + 1], P[3 * i + 2]);
    }
    memset(T->cstart, 0, sizeof(int32_t) * ((size_t)H->nkeys + 1));
    for (int i = 0; i < n; ++i) {
        ++T->cstart[T->key[i] + 1];
    }
    for (int k = 0; k < H->nkeys; ++k) {
        T->cstart[k + 1] += T->cstart[k];
    }
    /* stable scatter: ascending id within a cell (src/v2m_particles.cpp:791-795) */
    int32_t* cur = malloc(sizeof(int32_t) * ((size_t)H->nkeys + 1));
    memcpy(cur, T->cstart, sizeof(int32_t) * ((size_t)H->nkeys + 1));
    for (int i = 0; i < n; ++i) {
        T->sorted[cur[T->key[i]]++] = i;
    }
    free(cur);
#pragma omp parallel for schedule(static)
    for (int s = 0; s < n; ++s) {
        const int j = T->sorted[s];
        T->Ps[4 * (size_t)s] = P[3 * j];
        T->Ps[4 * (size_t)s + 1] = P[3 * j + 1];
        T->Ps[4 * (size_t)s + 2] = P[3 * j + 2];
        T->Ps[4 * (size_t)s + 3] = hn[j];
    }
    NbrIn in = { hn, P, lab, typ, T->cstart, T->sorted, T->Ps, prm->t, prm->skin };
#pragma omp parallel for schedule(dynamic, 256)
    for (int i = 0; i < n; ++i) {
        truss_of(H, &in, i, T->nbr, T->nnb);
    }
}

int orc_rebuild(const tn_grid_view* g, const tn_relax_params* prm, const tn_nodes_view* nd, int32_t* key,
                int32_t* sorted, int32_t* cstart, int32_t* nbr, int32_t* nnb, float* hn) {
    if (supported(g, prm)) {
        return 1;
    }
    const Fld f = fld_of(g);
    const int n = (int)nd->n;
    Truss T;
    make_hash(g, prm, &T.H);
    if (truss_alloc(&T, n)) {
        return 1;
    }
    for (int i = 0; i < n; ++i) {
        hn[i] = size_at(&f, nd->P[3 * i], nd->P[3 * i + 1], nd->P[3 * i + 2]);
    }
    truss_rebuild(&T, &f, prm, nd->P, nd->lab, nd->typ, hn);
    if (key) {
        memcpy(key, T.key, sizeof(int32_t) * (size_t)n);
    }
    if (sorted) {
        memcpy(sorted, T.sorted, sizeof(int32_t) * (size_t)n);
    }
    if (cstart) {
        memcpy(cstart, T.cstart, sizeof(int32_t) * ((size_t)T.H.nkeys + 1));
    }
    if (nbr) {
        memcpy(nbr, T.nbr, sizeof(int32_t) * (size_t)n * TN_K);
    }
    if (nnb) {
        memcpy(nnb, T.nnb, sizeof(int32_t) * (size_t)n);
    }
    truss_free(&T);
    return 0;
}

void orc_force(const tn_relax_params* prm, const tn_nodes_view* nd, const float* hn, const int32_t* nbr,
               const int32_t* nnb, float* F) {
    const int n = (int)nd->n;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
        bar_force(hn, nd->P, nd->typ, nbr, nnb, prm->fscale, prm->fsurf, i, &F[4 * (size_t)i]);
    }
}

int orc_move(const tn_grid_view* g, const tn_relax_params* prm, tn_nodes_view* nd, float* hn, const float* F,
             float* V, float fdt, float falpha, float* pw, float* mv) {
    if (supported(g, prm)) {
        return 1;
    }
    const Fld f = fld_of(g);
    const int n = (int)nd->n;
    const int fire = prm->fire && !prm->voxel_trap ? 1 : 0;
    NodeIO io = { nd->P, nd->lab, nd->typ, nd->part, hn };
#pragma omp parallel for schedule(dynamic, 1024)
    for (int i = 0; i < n; ++i) {
        mv[i] = node_move(&f, F, prm->dt, prm->maxstep, prm->snap, prm->voxel_trap != 0, i, &io, fire, V, fdt, falpha,
                          pw);
    }
    return 0;
}

int orc_relax(const tn_grid_view* g, const tn_relax_params* prm, tn_nodes_view* nd, tn_relax_stats* st,
              orc_iter_trace* trace, int32_t nthreads) {
    if (supported(g, prm) || !nd || !st) {
        return 1;
    }
#ifdef _OPENMP
    const int nthr0 = omp_get_max_threads();
    if (nthreads > 0) {
        omp_set_num_threads(nthreads);
    }
#endif
    const int rebuild_div = prm->rebuild_div > 0 ? prm->rebuild_div : 300;
    const Fld f = fld_of(g);
    const int n = (int)nd->n;
    memset(st, 0, sizeof *st);
    Truss T;
    make_hash(g, prm, &T.H);
    if (truss_alloc(&T, n)) {
        return 1;
    }
    const size_t n1 = (size_t)(n > 0 ? n : 1);
    float* F = malloc(sizeof(float) * 4 * n1);
    float* P0 = malloc(sizeof(float) * 3 * n1);
    float* mv = malloc(sizeof(float) * n1);
    float* hn = malloc(sizeof(float) * n1);
    const int fire = prm->fire && !prm->voxel_trap ? 1 : 0;
    float* V = calloc(fire ? 3 * n1 : 1, sizeof(float));
    float* pw = calloc(fire ? n1 : 1, sizeof(float));
    if (!F || !P0 || !mv || !hn || !V || !pw) {
        free(F), free(P0), free(mv), free(hn), free(V), free(pw);
        truss_free(&T);
        return fail("oracle: out of memory");
    }
    float* P = nd->P;
    memcpy(P0, P, sizeof(float) * 3 * (size_t)n);
    for (int i = 0; i < n; ++i) {
        hn[i] = size_at(&f, P[3 * i], P[3 * i + 1], P[3 * i + 2]);
    }
    NodeIO io = { nd->P, nd->lab, nd->typ, nd->part, hn };

    /* rebuild: lists, then P0 = P */
#define REBUILD()                                                          \
    do {                                                                   \
        truss_rebuild(&T, &f, prm, P, nd->lab, nd->typ, hn);              \
        memcpy(P0, P, sizeof(float) * 3 * (size_t)n);                      \
        ++st->rebuilds;                                                    \
    } while (0)

    REBUILD();
    Fire fc;
    fire_init(&fc, prm);
    for (int it = 0; it < prm->max_iters; ++it) {
#pragma omp parallel for schedule(dynamic, 1024)
        for (int i = 0; i < n; ++i) {
            bar_force(hn, P, nd->typ, T.nbr, T.nnb, prm->fscale, prm->fsurf, i, &F[4 * (size_t)i]);
        }
        float mmax = 0.0f;
#pragma omp parallel for schedule(dynamic, 1024) reduction(max : mmax)
        for (int i = 0; i < n; ++i) {
            mv[i] = node_move(&f, F, prm->dt, prm->maxstep, prm->snap, prm->voxel_trap != 0, i, &io, fire, V, fc.dt,
                              fc.alpha, pw);
            mmax = mmax > mv[i] ? mmax : mv[i];
        }
        double pt = 0.0;
        if (fire) {
            for (int i = 0; i < n; ++i) { /* serial, index order */
                pt += pw[i];
            }
            if (fire_update(&fc, pt)) {
                memset(V, 0, sizeof(float) * 3 * (size_t)n);
            }
        }
        st->iters = it + 1;
        st->last_move = mmax;
        st->fire_resets = fc.resets;
        st->fire_dt = fire ? fc.dt : 0.0f;

        int hist[32] = { 0 };
        for (int i = 0; i < n; ++i) {
            ++hist[orc_hist_bin(mv[i])];
        }
        int acc = 0, b99 = 31;
        for (int b = 0; b < 32; ++b) {
            acc += hist[b];
            if (acc >= 0.99 * n) {
                b99 = b;
                break;
            }
        }
        const float p99 = ldexpf(1.0f, b99 - 20 + 1);
        st->last_p99 = p99;
        orc_iter_trace* tr = trace ? &trace[it] : NULL;
        if (tr) {
            memset(tr, 0, sizeof *tr);
            tr->it = it;
            tr->mmax = mmax;
            tr->p99 = p99;
            tr->fdt = fc.dt;
            tr->falpha = fc.alpha;
            tr->power = pt;
            memcpy(tr->hist, hist, sizeof hist);
        }
        if (p99 < prm->dptol && (!fire || fire_may_stop(&fc))) {
            break;
        }

        /* Verlet bookkeeping, src/v2m_particles.cpp:917-937 */
        int nhalf = 0, nref = 0;
        NbrIn in = { hn, P, nd->lab, nd->typ, T.cstart, T.sorted, T.Ps, prm->t, prm->skin };
#pragma omp parallel for schedule(dynamic, 1024) reduction(+ : nhalf, nref)
        for (int i = 0; i < n; ++i) {
            const float ex = P[3 * i] - P0[3 * i], ey = P[3 * i + 1] - P0[3 * i + 1], ez = P[3 * i + 2] - P0[3 * i + 2];
            const float h = size_at(&f, P[3 * i], P[3 * i + 1], P[3 * i + 2]);
            const float m2 = ex * ex + ey * ey + ez * ez, s2 = prm->skin * prm->skin * h * h;
            if (m2 > s2) {
                truss_of(&T.H, &in, i, T.nbr, T.nnb);
                P0[3 * i] = P[3 * i];
                P0[3 * i + 1] = P[3 * i + 1];
                P0[3 * i + 2] = P[3 * i + 2];
                ++nref;
            } else {
                nhalf += m2 > 0.25f * s2;
            }
        }
        const int endgame = it >= prm->max_iters * 4 / 5 || p99 < 4.0f * prm->dptol;
        const int div = endgame ? (rebuild_div > 1000 ? rebuild_div : 1000) : rebuild_div;
        if (tr) {
            tr->nhalf = nhalf;
            tr->nrefresh = nref;
        }
        if (nhalf > n / div) {
            REBUILD();
            if (tr) {
                tr->rebuilt = 1;
            }
        }
    }
#undef REBUILD
    st->n_interior = st->n_interface = st->n_junction = st->n_corner = 0;
    for (int i = 0; i < n; ++i) {
        const int t = nd->typ[i];
        st->n_interior += t == TN_INTERIOR;
        st->n_interface += t == TN_INTERFACE;
        st->n_junction += t == TN_JUNCTION;
        st->n_corner += t == TN_CORNER;
    }
    free(F), free(P0), free(mv), free(hn), free(V), free(pw);
    truss_free(&T);
#ifdef _OPENMP
    if (nthreads > 0) {
        omp_set_num_threads(nthr0);
    }
#endif
    return 0;
}

void orc_field_points(const tn_grid_view* g, int32_t a, int32_t b, int64_t m, const float* X, float* phi_a,
                      float* psi, float* grad, int32_t* label, int32_t* second) {
    const Fld f = fld_of(g);
    for (int64_t k = 0; k < m; ++k) {
        const float* x = X + 3 * k;
        phi_a[k] = phi_point(&f, a, x[0], x[1], x[2]);
        psi[k] = psi_grad_point(&f, a, b, x[0], x[1], x[2], grad + 3 * k);
        int sec;
        float mg;
        label[k] = owner_at(&f, x[0], x[1], x[2], &sec, &mg);
        second[k] = sec;
    }
}

void orc_sdf_points(const float* prog, int32_t l, int64_t m, const float* X, int32_t grad, float* v, float* g) {
    for (int64_t k = 0; k < m; ++k) {
        const SdfV r = sdf_evalv(prog, l, X[3 * k], X[3 * k + 1], X[3 * k + 2], grad);
        v[k] = r.v;
        if (g) {
            g[3 * k] = r.x;
            g[3 * k + 1] = r.y;
            g[3 * k + 2] = r.z;
        }
    }
}
