/*
 * thin.c -- CPU restatement of v2mesh's node thinning and feature protection
 * (see oracle.h): thin_nodes (src/v2m_particles.cpp:1223-1357) and
 * protect_features (src/v2m_particles.cpp:1002-1221).
 *
 * TEST INFRASTRUCTURE ONLY: the checker of the HIP path, never linked by it.
 * Both functions return a keep flag per node; the reference then compacts the
 * kept nodes in index order. The arithmetic keeps the reference's operations
 * and their order (build with -ffp-contract=off).
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* v2m_h_at, src/opencl/v2m_particle_body.cl:30-36 */
static float h_at(const tn_grid_view* g, float px, float py, float pz) {
    int i = (int)floorf(px / g->vs[0] + 0.5f), j = (int)floorf(py / g->vs[1] + 0.5f),
        k = (int)floorf(pz / g->vs[2] + 0.5f);
    i = i < 0 ? 0 : (i >= g->nx ? g->nx - 1 : i);
    j = j < 0 ? 0 : (j >= g->ny ? g->ny - 1 : j);
    k = k < 0 ? 0 : (k >= g->nz ? g->nz - 1 : k);
    return g->h[i + (size_t)g->nx * (j + (size_t)g->ny * k)];
}

/* ---- thin_nodes ------------------------------------------------------------------ */

/* the cell key of src/v2m_particles.cpp:1248-1255 */
static uint64_t thin_cell(int L, float c0, float x, float y, float z, int dx, int dy, int dz) {
    const float c = c0 * (float)(1 << L);
    const uint64_t ix = (uint64_t)((int64_t)floorf(x / c) + dx + (1 << 19)) & 0xfffff,
                   iy = (uint64_t)((int64_t)floorf(y / c) + dy + (1 << 19)) & 0xfffff,
                   iz = (uint64_t)((int64_t)floorf(z / c) + dz + (1 << 19)) & 0xfffff;
    return (uint64_t)L << 60 | ix << 40 | iy << 20 | iz;
}

/* an open-addressing map from cell keys to lists of nodes (the reference's
 * unordered_map<uint64_t, vector<int>>; list order does not matter here) */
typedef struct {
    uint64_t* key;
    int32_t* head;
    int64_t cap;
    int32_t* next;
    int32_t* node;
    int64_t used;
} CellMap;

static uint64_t mix64(uint64_t k) {
    k ^= k >> 33;
    k *= 0xff51afd7ed558ccdULL;
    k ^= k >> 33;
    return k;
}

static int64_t map_find(const CellMap* m, uint64_t key) {
    int64_t s = (int64_t)(mix64(key) & (uint64_t)(m->cap - 1));
    while (m->head[s] != -1) {
        if (m->key[s] == key) {
            return s;
        }
        s = (s + 1) & (m->cap - 1);
    }
    return -1 - s;   /* not found: -1 - (the empty slot) */
}

static void map_add(CellMap* m, uint64_t key, int32_t j) {
    int64_t s = map_find(m, key);
    if (s < 0) {
        s = -1 - s;
        m->key[s] = key;
        m->head[s] = -2;   /* present, empty list */
    }
    m->node[m->used] = j;
    m->next[m->used] = m->head[s] == -2 ? -1 : m->head[s];
    m->head[s] = (int32_t)m->used;
    ++m->used;
}

typedef struct {
    int ta;
    float hn;
    int32_t i;
} ThinCand;

static int thin_cmp(const void* pa, const void* pb) {   /* interfaces first, then h, then index */
    const ThinCand* a = (const ThinCand*)pa;
    const ThinCand* b = (const ThinCand*)pb;
    if (a->ta != b->ta) {
        return a->ta < b->ta ? -1 : 1;
    }
    if (a->hn != b->hn) {
        return a->hn < b->hn ? -1 : 1;
    }
    return a->i < b->i ? -1 : (a->i > b->i ? 1 : 0);
}

/* label l among node j's labels (the `has` of src/v2m_particles.cpp:1267-1270) */
static int has_label(const uint16_t* lab, const uint8_t* typ, const uint16_t* part, const uint16_t* part3, int j,
                     int l) {
    return lab[j] == l || (typ[j] != TN_INTERIOR && part[2 * j] == l) ||
           (typ[j] >= TN_JUNCTION && part[2 * j + 1] == l) ||
           (typ[j] == TN_CORNER && (part3 ? part3[j] : TN_NOLAB) == l);
}

int64_t orc_thin_nodes(const tn_grid_view* g, float B, int64_t n64, const float* P, const uint16_t* lab,
                       const uint8_t* typ, const uint16_t* part, const uint16_t* part3, uint8_t* keep) {
    const int n = (int)n64;
    if (B <= 0.0f || n == 0) {
        for (int i = 0; i < n; ++i) {
            keep[i] = 1;
        }
        return 0;
    }
    float* hn = (float*)malloc(sizeof(float) * (size_t)(n > 0 ? n : 1));
    for (int i = 0; i < n; ++i) {
        hn[i] = h_at(g, P[3 * i], P[3 * i + 1], P[3 * i + 2]);
    }
    const float c0 = fmaxf(1e-6f, B * g->hmin);
    int nlev = 1;
    while (nlev < 12 && c0 * (float)(1 << (nlev - 1)) < B * g->hmax) {
        ++nlev;
    }
    CellMap m;
    m.cap = 1;
    while (m.cap < 2 * (int64_t)n * nlev + 16) {
        m.cap <<= 1;
    }
    m.key = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)m.cap);
    m.head = (int32_t*)malloc(sizeof(int32_t) * (size_t)m.cap);
    for (int64_t s = 0; s < m.cap; ++s) {
        m.head[s] = -1;
    }
    m.next = (int32_t*)malloc(sizeof(int32_t) * (size_t)n * nlev);
    m.node = (int32_t*)malloc(sizeof(int32_t) * (size_t)n * nlev);
    m.used = 0;
    memset(keep, 0, (size_t)n);
#define THIN_INSERT(j)                                                                               \
    do {                                                                                             \
        keep[j] = 1;                                                                                 \
        for (int L_ = 0; L_ < nlev; ++L_) {                                                          \
            map_add(&m, thin_cell(L_, c0, P[3 * (j)], P[3 * (j) + 1], P[3 * (j) + 2], 0, 0, 0), (j)); \
        }                                                                                            \
    } while (0)
    for (int j = 0; j < n; ++j) {   /* junction / corner nodes: always kept */
        if (typ[j] >= TN_JUNCTION) {
            THIN_INSERT(j);
        }
    }
    ThinCand* cand = (ThinCand*)malloc(sizeof(ThinCand) * (size_t)(n > 0 ? n : 1));
    int nc = 0;
    for (int i = 0; i < n; ++i) {
        if (typ[i] == TN_INTERFACE || typ[i] == TN_INTERIOR) {
            cand[nc].ta = typ[i] == TN_INTERIOR;
            cand[nc].hn = hn[i];
            cand[nc].i = i;
            ++nc;
        }
    }
    qsort(cand, (size_t)nc, sizeof(ThinCand), thin_cmp);
    int64_t removed = 0;
    for (int c = 0; c < nc; ++c) {
        const int i = cand[c].i;
        const float r = B * hn[i], r2 = r * r;
        int L = 0;
        while (L + 1 < nlev && c0 * (float)(1 << L) < r) {
            ++L;
        }
        const float x = P[3 * i], y = P[3 * i + 1], z = P[3 * i + 2];
        const int iface = typ[i] == TN_INTERFACE;
        const int a = lab[i], b = iface ? part[2 * i] : -1;
        int crowded = 0;
        for (int dz = -1; dz <= 1 && !crowded; ++dz) {
            for (int dy = -1; dy <= 1 && !crowded; ++dy) {
                for (int dx = -1; dx <= 1 && !crowded; ++dx) {
                    const int64_t s = map_find(&m, thin_cell(L, c0, x, y, z, dx, dy, dz));
                    if (s < 0) {
                        continue;
                    }
                    for (int32_t e = m.head[s]; e >= 0; e = m.next[e]) {
                        const int j = m.node[e];
                        const float ex = P[3 * j] - x, ey = P[3 * j + 1] - y, ez = P[3 * j + 2] - z;
                        if (ex * ex + ey * ey + ez * ez >= r2) {
                            continue;
                        }
                        if (iface ? (typ[j] != TN_INTERIOR && has_label(lab, typ, part, part3, j, a) &&
                                     has_label(lab, typ, part, part3, j, b))
                                  : has_label(lab, typ, part, part3, j, a)) {
                            crowded = 1;
                            break;
                        }
                    }
                }
            }
        }
        if (crowded) {
            ++removed;
        } else {
            THIN_INSERT(i);
        }
    }
#undef THIN_INSERT
    free(cand);
    free(m.key);
    free(m.head);
    free(m.next);
    free(m.node);
    free(hn);
    return removed;
}

/* ---- protect_features ------------------------------------------------------------ */

/* the distance of p from the segment a-b and the parameter of its foot
 * (the `seg` of src/v2m_particles.cpp:1020-1038) */
static void seg_locate(const float* a, const float* b, const float* p, float* dd, float* tt) {
    float ab[3], ap[3], ab2 = 0, apab = 0;
    for (int k = 0; k < 3; ++k) {
        ab[k] = b[k] - a[k];
        ap[k] = p[k] - a[k];
        ab2 += ab[k] * ab[k];
        apab += ap[k] * ab[k];
    }
    *tt = ab2 > 0 ? fminf(1.0f, fmaxf(0.0f, apab / ab2)) : 0.0f;
    float d = 0;
    for (int k = 0; k < 3; ++k) {
        const float e = ap[k] - *tt * ab[k];
        d += e * e;
    }
    *dd = sqrtf(d);
}

/* locate, src/v2m_particles.cpp:1018-1084 */
static void feat_locate(int type, const float* q, const float* p, float* dist, float* t) {
    if (type == 2) {
        seg_locate(q, q + 3, p, dist, t);
    } else if (type == 4) {
        const int m = (int)q[0];
        *dist = 1e30f;
        for (int i = 0; i + 1 < m; ++i) {
            float dd, tt;
            seg_locate(q + 1 + 3 * i, q + 4 + 3 * i, p, &dd, &tt);
            if (dd < *dist) {
                *dist = dd;
                *t = (float)i + tt;
            }
        }
    } else {
        const float* c = q;
        const float* nn = q + 3;
        float v[3], h = 0;
        for (int k = 0; k < 3; ++k) {
            v[k] = p[k] - c[k];
            h += v[k] * nn[k];
        }
        float rad[3], rho = 0;
        for (int k = 0; k < 3; ++k) {
            rad[k] = v[k] - h * nn[k];
            rho += rad[k] * rad[k];
        }
        rho = sqrtf(rho);
        *dist = hypotf(h, rho - q[6]);
        const float a0 = fabsf(nn[0]) < 0.9f ? 1.0f : 0.0f, a1 = a0 > 0 ? 0.0f : 1.0f;
        float u[3] = { -nn[2] * a1, nn[2] * a0, nn[0] * a1 - nn[1] * a0 };
        const float lu = sqrtf(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
        for (int k = 0; k < 3; ++k) {
            u[k] /= lu;
        }
        const float w[3] = { nn[1] * u[2] - nn[2] * u[1], nn[2] * u[0] - nn[0] * u[2], nn[0] * u[1] - nn[1] * u[0] };
        *t = atan2f(rad[0] * w[0] + rad[1] * w[1] + rad[2] * w[2], rad[0] * u[0] + rad[1] * u[1] + rad[2] * u[2]);
    }
}

typedef struct {
    float t;
    int32_t j;
} PinAt;

static int pin_cmp(const void* pa, const void* pb) {   /* std::pair<float, int> order */
    const PinAt* a = (const PinAt*)pa;
    const PinAt* b = (const PinAt*)pb;
    if (a->t != b->t) {
        return a->t < b->t ? -1 : 1;
    }
    return a->j < b->j ? -1 : (a->j > b->j ? 1 : 0);
}

typedef struct {
    uint64_t key;
    int32_t ball;
} BallCell;

static int ballcell_cmp(const void* pa, const void* pb) {
    const BallCell* a = (const BallCell*)pa;
    const BallCell* b = (const BallCell*)pb;
    if (a->key != b->key) {
        return a->key < b->key ? -1 : 1;
    }
    return a->ball < b->ball ? -1 : (a->ball > b->ball ? 1 : 0);
}

static uint64_t ball_key(int64_t ix, int64_t iy, int64_t iz) {   /* src/v2m_particles.cpp:1131-1134 */
    return (uint64_t)(ix + (1 << 20)) << 42 | (uint64_t)(iy + (1 << 20)) << 21 | (uint64_t)(iz + (1 << 20));
}

int64_t orc_protect_features(const tn_grid_view* g, const float* feat, int64_t nfeat, int64_t n64, const float* P,
                             const uint8_t* typ, uint8_t* keep) {
    const int64_t n = n64;
    for (int64_t i = 0; i < n; ++i) {
        keep[i] = 1;
    }
    if (nfeat == 0 || n == 0) {
        return 0;
    }
    int32_t* pins = (int32_t*)malloc(sizeof(int32_t) * (size_t)n);
    int npin = 0;
    for (int64_t i = 0; i < n; ++i) {
        if (typ[i] == TN_CORNER) {
            pins[npin++] = (int32_t)i;
        }
    }
    float* balls = (float*)malloc(sizeof(float) * 4 * (size_t)(npin + 2) * 2);
    int64_t nball = 0, ballcap = (int64_t)(npin + 2) * 2;
    PinAt* on = (PinAt*)malloc(sizeof(PinAt) * (size_t)(npin + 1));
    for (int64_t k = 0; k < nfeat;) {
        const int type = (int)feat[k];
        const float* q = &feat[k + 1];
        k += type == 1 ? 4 : type == 2 ? 7 : type == 4 ? 2 + 3 * (int64_t)q[0] : 8;
        if (type == 1 || (type == 4 && q[0] < 2)) {
            continue;
        }
        int m = 0;
        for (int pi = 0; pi < npin; ++pi) {
            const int j = pins[pi];
            const float* p = &P[3 * (size_t)j];
            float dist, t = 0;
            feat_locate(type, q, p, &dist, &t);
            if (dist < 1e-3f * h_at(g, p[0], p[1], p[2])) {
                on[m].t = t;
                on[m].j = j;
                ++m;
            }
        }
        qsort(on, (size_t)m, sizeof(PinAt), pin_cmp);
        for (int i = 0; i + 1 < m + (type == 3 && m > 2 ? 1 : 0); ++i) {
            const float* a = &P[3 * (size_t)on[i].j];
            const float* b = &P[3 * (size_t)on[(i + 1) % m].j];
            const float c[3] = { 0.5f * (a[0] + b[0]), 0.5f * (a[1] + b[1]), 0.5f * (a[2] + b[2]) };
            const float r2 = 0.25f * ((a[0] - b[0]) * (a[0] - b[0]) + (a[1] - b[1]) * (a[1] - b[1]) +
                                      (a[2] - b[2]) * (a[2] - b[2]));
            const float h = h_at(g, c[0], c[1], c[2]);
            if (r2 > 0 && r2 <= 0.64f * h * h) {
                if (nball == ballcap) {
                    ballcap *= 2;
                    balls = (float*)realloc(balls, sizeof(float) * 4 * (size_t)ballcap);
                }
                balls[4 * nball] = c[0];
                balls[4 * nball + 1] = c[1];
                balls[4 * nball + 2] = c[2];
                balls[4 * nball + 3] = r2;
                ++nball;
            }
        }
    }
    free(on);
    free(pins);
    if (nball == 0) {
        free(balls);
        return 0;
    }
    float cs = 0;
    for (int64_t bi = 0; bi < nball; ++bi) {
        cs = fmaxf(cs, 2.0f * sqrtf(balls[4 * bi + 3]));
    }
    int64_t ncell = 0, cellcap = 64;
    BallCell* cells = (BallCell*)malloc(sizeof(BallCell) * (size_t)cellcap);
    for (int64_t bi = 0; bi < nball; ++bi) {
        const float* bl = &balls[4 * bi];
        const float r = sqrtf(bl[3]);
        int64_t lo[3], hi[3];
        for (int k = 0; k < 3; ++k) {
            lo[k] = (int64_t)floorf((bl[k] - r) / cs);
            hi[k] = (int64_t)floorf((bl[k] + r) / cs);
        }
        for (int64_t x = lo[0]; x <= hi[0]; ++x) {
            for (int64_t y = lo[1]; y <= hi[1]; ++y) {
                for (int64_t z = lo[2]; z <= hi[2]; ++z) {
                    if (ncell == cellcap) {
                        cellcap *= 2;
                        cells = (BallCell*)realloc(cells, sizeof(BallCell) * (size_t)cellcap);
                    }
                    cells[ncell].key = ball_key(x, y, z);
                    cells[ncell].ball = (int32_t)bi;
                    ++ncell;
                }
            }
        }
    }
    qsort(cells, (size_t)ncell, sizeof(BallCell), ballcell_cmp);
    int64_t dropped = 0;
    for (int64_t i = 0; i < n; ++i) {
        if (typ[i] == TN_CORNER) {
            continue;
        }
        const float* p = &P[3 * (size_t)i];
        const uint64_t key = ball_key((int64_t)floorf(p[0] / cs), (int64_t)floorf(p[1] / cs), (int64_t)floorf(p[2] / cs));
        int64_t lo = 0, hi = ncell;   /* the first entry with this key */
        while (lo < hi) {
            const int64_t mid = (lo + hi) / 2;
            if (cells[mid].key < key) {
                lo = mid + 1;
            } else {
                hi = mid;
            }
        }
        for (int64_t e = lo; e < ncell && cells[e].key == key; ++e) {
            const float* bl = &balls[4 * (int64_t)cells[e].ball];
            const float dx = p[0] - bl[0], dy = p[1] - bl[1], dz = p[2] - bl[2];
            if (dx * dx + dy * dy + dz * dz < bl[3]) {
                keep[i] = 0;
                ++dropped;
                break;
            }
        }
    }
    free(cells);
    free(balls);
    return dropped;
}
