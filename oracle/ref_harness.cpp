// ref_harness.cpp -- a C entry layer over the reference's own CPU relaxation,
// compiled from the reference sources where they lie (oracle/Makefile, target
// `ref`) into oracle/_ref/libv2m_ref.so. TEST INFRASTRUCTURE ONLY: it builds
// golden fixtures (tests/golden/make_golden.py) and times the reference CPU
// path as the CPU baseline; nothing in the product links it.
//
// What it calls:
//   build_grid_cpu  src/v2m_grid.cpp:47-722        (the field the nodes move in)
//   seed_cpu        src/v2m_particles.cpp:625-635  (the graded HCP seeds)
//   relax_cpu       src/v2m_particles.cpp:728-976  (the relaxation itself)
// and, for stage-by-stage fixtures, the per-node bodies the reference includes
// into its drivers (src/v2m_particles.cpp:26-46 includes them the same way):
//   v2m_h_at / v2m_level_of / v2m_bin_key / v2m_neighbors / v2m_force / v2m_move
//   (src/opencl/v2m_particle_body.cl).

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "delaunay.h"
#include "v2m_gdel.h"
#include "v2m_grid.h"
#include "v2m_particles.h"
#include "v2m_tetra.h"
#include "v2m_sdfshape.h"
#include "v2m_tpm.h"
#include "v2m_volume.h"

#include "../include/trussnet_amd.h"

namespace refh {
using std::ceil;
using std::exp;
using std::fabs;
using std::floor;
using std::fmax;
using std::fmin;
using std::log;
using std::pow;
using std::sqrt;
typedef uint8_t uchar;
typedef uint16_t ushort;
#define V2M_G
#include "opencl/v2m_grid_body.cl"
#include "opencl/v2m_sdf_body.cl"
#include "opencl/v2m_seed_body.cl"
#include "opencl/v2m_particle_body.cl"
#undef V2M_G
}  // namespace refh

namespace {

struct RefGrid {
    tn::LabelVolume lv;
    tn::Grid g;
    tn::Nodes seeds;
};

thread_local std::string g_err;

refh::V2mDims dims_of(const tn::Grid& g) {
    refh::V2mDims d;
    d.nx = g.nx;
    d.ny = g.ny;
    d.nz = g.nz;
    d.nbx = g.nbx;
    d.nby = g.nby;
    d.nbz = g.nbz;
    d.vx = g.vs[0];
    d.vy = g.vs[1];
    d.vz = g.vs[2];
    return d;
}

tn::RelaxParams to_params(const tn_relax_params* p) {
    tn::RelaxParams r;
    r.nseed = p->nseed;
    r.t = p->t;
    r.skin = p->skin;
    r.fscale = p->fscale;
    r.fsurf = p->fsurf;
    r.dt = p->dt;
    r.maxstep = p->maxstep;
    r.snap = p->snap;
    r.max_iters = p->max_iters;
    r.dptol = p->dptol;
    r.jseed = p->jseed;
    r.corners = p->corners != 0;
    r.voxel_trap = p->voxel_trap != 0;
    r.thin = p->thin;
    r.fire = p->fire != 0;
    r.fire_dtmax = p->fire_dtmax;
    r.verbose = p->verbose != 0;
    return r;
}

// the hash geometry exactly as relax_cpu sets it up (src/v2m_particles.cpp:738-761)
refh::V2mHash hash_of(const tn::Grid& g, const tn::RelaxParams& prm, int* nkeys_out) {
    refh::V2mHash H;
    std::memset(&H, 0, sizeof H);
    H.b0 = (prm.t + prm.skin) * g.hmin;
    H.ox = -0.5f * g.vs[0];
    H.oy = -0.5f * g.vs[1];
    H.oz = -0.5f * g.vs[2];
    const float ext[3] = { g.nx * g.vs[0], g.ny * g.vs[1], g.nz * g.vs[2] };
    H.nlev = 1;

    while (H.nlev < V2M_MAXLEV && H.b0 * static_cast<float>(1 << (H.nlev - 1)) < (prm.t + prm.skin) * g.hmax) {
        ++H.nlev;
    }

    int nkeys = 0;

    for (int L = 0; L < H.nlev; ++L) {
        H.off[L] = nkeys;
        const float bL = H.b0 * static_cast<float>(1 << L);

        for (int a = 0; a < 3; ++a) {
            H.dim[L][a] = std::max(1, static_cast<int>(std::ceil(ext[a] / bL)));
        }

        nkeys += H.dim[L][0] * H.dim[L][1] * H.dim[L][2];
    }

    *nkeys_out = nkeys;
    return H;
}

}  // namespace

extern "C" {

typedef struct ref_grid_params {
    float sigma, hbase, hmin, hmax, K, g;
} ref_grid_params;

const char* ref_last_error(void) {
    return g_err.c_str();
}

void* ref_grid_build_prob(const uint16_t* labels, const float* prob, int nprob, int nx, int ny, int nz,
                          const float* vs, const ref_grid_params* gp, const float* hvox);
void* ref_grid_build_gray(const uint16_t* labels, const float* gray, const float* thr, int m, int nx, int ny, int nz,
                          const float* vs, const ref_grid_params* gp, const float* hvox);
void* ref_grid_build_gray_prob(const uint16_t* labels, const float* prob, int nprob, const float* gray,
                               const float* thr, int m_thr, int nx, int ny, int nz, const float* vs,
                               const ref_grid_params* gp, const float* hvox);

// build_grid_cpu over a label volume (x fastest); hvox may be null
void* ref_grid_build(const uint16_t* labels, int nx, int ny, int nz, const float* vs, const ref_grid_params* gp,
                     const float* hvox) {
    return ref_grid_build_prob(labels, nullptr, 0, nx, ny, nz, vs, gp, hvox);
}

// the same with per-label probabilities (a tissue-probability input, label-major
// prob[l * nv + v], label 0 the exterior): the interface fields are then the
// smoothed probabilities (src/v2m_grid.cpp:154-168, 467)
void* ref_grid_build_prob(const uint16_t* labels, const float* prob, int nprob, int nx, int ny, int nz,
                          const float* vs, const ref_grid_params* gp, const float* hvox) {
    return ref_grid_build_gray_prob(labels, prob, nprob, nullptr, nullptr, 0, nx, ny, nz, vs, gp, hvox);
}

// the same over a gray-scale volume: gray[v] the intensity, thr the m ascending
// thresholds, labels[v] the number of thresholds <= gray[v] (apply_thresholds with
// no smoothing, src/v2m_volume.cpp:93-152); the fields are then the memberships of
// the intensity (src/v2m_grid.cpp:374-433)
void* ref_grid_build_gray(const uint16_t* labels, const float* gray, const float* thr, int m, int nx, int ny, int nz,
                          const float* vs, const ref_grid_params* gp, const float* hvox) {
    return ref_grid_build_gray_prob(labels, nullptr, 0, gray, thr, m, nx, ny, nz, vs, gp, hvox);
}

void* ref_grid_build_gray_prob(const uint16_t* labels, const float* prob, int nprob, const float* gray,
                               const float* thr, int m_thr, int nx, int ny, int nz, const float* vs,
                               const ref_grid_params* gp, const float* hvox) {
    try {
        RefGrid* r = new RefGrid;
        const size_t nv = static_cast<size_t>(nx) * ny * nz;
        r->lv.data.assign(labels, labels + nv);

        if (prob && nprob > 0) {
            r->lv.prob.assign(prob, prob + nv * nprob);
            r->lv.nprob = nprob;
        }

        if (gray && thr && m_thr > 0) {
            r->lv.gray.assign(gray, gray + nv);
            r->lv.thresholds.assign(thr, thr + m_thr);
        }

        r->lv.nx = nx;
        r->lv.ny = ny;
        r->lv.nz = nz;

        for (int a = 0; a < 3; ++a) {
            r->lv.voxelsize[a] = vs[a];
        }

        int m = 0;

        for (uint16_t l : r->lv.data) {
            m = std::max<int>(m, l);
        }

        r->lv.maxlabel = m;
        tn::GridParams p;
        p.sigma = gp->sigma;
        p.hbase = gp->hbase;
        p.hmin = gp->hmin;
        p.hmax = gp->hmax;
        p.K = gp->K;
        p.g = gp->g;

        if (hvox) {
            p.hvox.assign(hvox, hvox + nv);
        }

        tn::build_grid_cpu(r->lv, p, r->g);
        return r;
    } catch (const std::exception& e) {
        g_err = e.what();
        return nullptr;
    }
}

// build_grid_cpu with every GridParams field, from the device build's own views
// (a label volume, a probability input or a gray-scale input)
void* ref_grid_build_full(const tn_volume_view* v, const tn_grid_params* gp) {
    try {
        RefGrid* r = new RefGrid;
        const size_t nv = static_cast<size_t>(v->nx) * v->ny * v->nz;
        r->lv.data.assign(v->L, v->L + nv);
        r->lv.nx = v->nx;
        r->lv.ny = v->ny;
        r->lv.nz = v->nz;

        for (int a = 0; a < 3; ++a) {
            r->lv.voxelsize[a] = v->vs[a];
        }

        if (v->prob && v->nprob > 0) {
            r->lv.prob.assign(v->prob, v->prob + nv * v->nprob);
            r->lv.nprob = v->nprob;
        }

        if (v->gray && v->thresholds && v->nthr > 0) {
            r->lv.gray.assign(v->gray, v->gray + nv);
            r->lv.thresholds.assign(v->thresholds, v->thresholds + v->nthr);
        }

        int m = 0;

        for (uint16_t l : r->lv.data) {
            m = std::max<int>(m, l);
        }

        r->lv.maxlabel = m;
        tn::GridParams p;
        p.sigma = gp->sigma;
        p.sigma_thin = gp->sigma_thin;
        p.thin_lo = gp->thin_lo;
        p.thin_hi = gp->thin_hi;
        p.thick = gp->thick;
        p.thin_floor = gp->thin_floor;
        p.preserve = gp->preserve;
        p.sigma_curv = gp->sigma_curv;
        p.hbase = gp->hbase;
        p.hmin = gp->hmin;
        p.hmax = gp->hmax;
        p.K = gp->K;
        p.g = gp->g;

        if (gp->nhlab > 0) {
            p.hlab.assign(gp->hlab, gp->hlab + gp->nhlab);
        }

        p.isize.glob = gp->iglob;

        if (gp->nilab > 0) {
            p.isize.lab.assign(gp->ilab, gp->ilab + gp->nilab);
        }

        if (gp->npair > 0) {
            p.isize.pair.assign(gp->ipair, gp->ipair + 3 * static_cast<size_t>(gp->npair));
        }

        if (v->hvox) {
            p.hvox.assign(v->hvox, v->hvox + nv);
        }

        tn::build_grid_cpu(r->lv, p, r->g);
        return r;
    } catch (const std::exception& e) {
        g_err = e.what();
        return nullptr;
    }
}

// A grid from shape constructs (JSON text): the steps of shapes_volume
// (src/v2m_modes.cpp:508-545) -- load_shapes, the gap-closing radius and cull
// margin written into the program, the per-brick programs, the raster's fields
// through apply_tpm -- then build_grid_cpu with the analytic fields (gm = -1).
// voxel <= 0: hbase / 3 (or the scene's extent / 160).
void* ref_grid_build_shapes(const char* json_text, double voxel, const tn_grid_params* gp) {
    try {
        RefGrid* r = new RefGrid;
        tn::ShapeScene sc = tn::load_shapes(json_text, true, "overwrite");

        if (!(voxel > 0)) {
            const double ext = std::max(sc.hi[0] - sc.lo[0], std::max(sc.hi[1] - sc.lo[1], sc.hi[2] - sc.lo[2]));
            voxel = gp->hbase > 0 ? gp->hbase / 3.0 : ext / 160.0;
        }

        const double hb = gp->hbase > 0 ? gp->hbase : 3.0 * voxel;
        const double gap = 0.5;   // PipelineOptions::shape_gap
        sc.prog[6 + static_cast<size_t>(sc.nlab)] = static_cast<float>(gap * hb);
        sc.prog[7 + static_cast<size_t>(sc.nlab)] = static_cast<float>(std::max(1.25 * hb, std::max(gap * hb, 3.0 * voxel)));
        tn::build_brick_programs(sc, voxel);
        const tn::Tpm tpm = tn::rasterize_scene(sc, voxel);
        tn::TpmOptions to;
        to.fields = true;
        to.fill_holes = false;
        to.exterior.assign(1, 0);
        to.map.clear();
        to.sigma = 0.0f;
        tn::apply_tpm(tpm, to, r->lv);
        r->lv.sdf = sc.prog;
        r->lv.sdf_feat = sc.feat;
        tn::GridParams p;
        p.sigma = gp->sigma;
        p.sigma_thin = gp->sigma_thin;
        p.thin_lo = gp->thin_lo;
        p.thin_hi = gp->thin_hi;
        p.thick = gp->thick;
        p.thin_floor = gp->thin_floor;
        p.preserve = gp->preserve;
        p.sigma_curv = gp->sigma_curv;
        p.hbase = gp->hbase;
        p.hmin = gp->hmin;
        p.hmax = gp->hmax;
        p.K = gp->K;
        p.g = gp->g;
        tn::build_grid_cpu(r->lv, p, r->g);
        return r;
    } catch (const std::exception& e) {
        g_err = e.what();
        return nullptr;
    }
}

// the Grid's scalars: info[0..3] = R, limit_sweeps, overflow_bricks, nslot;
// sc[0..3] = hbase, hmin, hmax, hcurv
void ref_grid_scalars(void* h, int64_t* info, float* sc) {
    const tn::Grid& g = static_cast<RefGrid*>(h)->g;
    info[0] = g.R;
    info[1] = g.limit_sweeps;
    info[2] = g.overflow_bricks;
    info[3] = static_cast<int64_t>(g.phi.size() / V2M_SLOT);
    sc[0] = g.hbase;
    sc[1] = g.hmin;
    sc[2] = g.hmax;
    sc[3] = g.hcurv;
}

// a grid from raw arrays (a fixture), for relaxing it with the reference
void* ref_grid_from_view(const tn_grid_view* v) {
    RefGrid* r = new RefGrid;
    const size_t nv = static_cast<size_t>(v->nx) * v->ny * v->nz;
    const size_t nb = static_cast<size_t>(v->nbx) * v->nby * v->nbz;
    r->lv.data.assign(v->L, v->L + nv);
    r->lv.nx = v->nx;
    r->lv.ny = v->ny;
    r->lv.nz = v->nz;
    tn::Grid& g = r->g;
    g.nx = v->nx;
    g.ny = v->ny;
    g.nz = v->nz;
    g.nbx = v->nbx;
    g.nby = v->nby;
    g.nbz = v->nbz;

    for (int a = 0; a < 3; ++a) {
        g.vs[a] = v->vs[a];
    }

    g.L = &r->lv.data;
    g.bl_cnt.assign(v->bl_cnt, v->bl_cnt + nb);
    g.bl_lab.assign(v->bl_lab, v->bl_lab + nb * V2M_BL);
    g.bl_slot.assign(v->bl_slot, v->bl_slot + nb);
    g.phi.assign(v->phi, v->phi + static_cast<size_t>(v->nslot) * V2M_SLOT);
    g.h.assign(v->h, v->h + nv);
    g.hmin = v->hmin;
    g.hmax = v->hmax;
    g.gm = 0;
    g.gTW.assign(1, 0.0f);

    if (v->gm > 0 && v->gI && v->gTW && v->ngTW >= 2 * static_cast<int64_t>(v->gm)) {
        r->lv.gray.assign(v->gI, v->gI + nv);
        g.gI = r->lv.gray.data();
        g.gm = v->gm;
        g.gTW.assign(v->gTW, v->gTW + 2 * v->gm);
        g.gray_w.assign(v->gTW + v->gm, v->gTW + 2 * v->gm);
    } else if (v->gm < 0 && v->gTW && v->ngTW > 0) {   // shape fields: the compiled program
        g.gm = v->gm;
        g.gTW.assign(v->gTW, v->gTW + v->ngTW);
    }

    if (v->feat && v->nfeat > 0) {   // the feature curves
        g.feat.assign(v->feat, v->feat + v->nfeat);
    }

    return r;
}

// the per-voxel grades (Grid::grade) that seed_cpu reads, for a grid made from a view
void ref_grid_set_grade(void* h, const uint8_t* grade) {
    RefGrid* r = static_cast<RefGrid*>(h);
    const size_t nv = static_cast<size_t>(r->g.nx) * r->g.ny * r->g.nz;
    r->g.grade.assign(grade, grade + nv);
}

void ref_grid_free(void* h) {
    delete static_cast<RefGrid*>(h);
}

// pointers into the handle's own arrays (valid until ref_grid_free); grade[nv] too
void ref_grid_view(void* h, tn_grid_view* v, const uint8_t** grade, float* hbase) {
    RefGrid* r = static_cast<RefGrid*>(h);
    const tn::Grid& g = r->g;
    std::memset(v, 0, sizeof *v);
    v->nx = g.nx;
    v->ny = g.ny;
    v->nz = g.nz;
    v->nbx = g.nbx;
    v->nby = g.nby;
    v->nbz = g.nbz;

    for (int a = 0; a < 3; ++a) {
        v->vs[a] = g.vs[a];
    }

    v->L = g.L->data();
    v->bl_cnt = g.bl_cnt.data();
    v->bl_lab = g.bl_lab.data();
    v->bl_slot = g.bl_slot.data();
    v->phi = g.phi.data();
    v->nslot = static_cast<int64_t>(g.phi.size() / V2M_SLOT);
    v->h = g.h.data();
    v->hmin = g.hmin;
    v->hmax = g.hmax;
    v->gm = g.gm;

    if (g.gm != 0) {
        v->gI = g.gm > 0 ? g.gI : nullptr;
        v->gTW = g.gTW.data();
        v->ngTW = static_cast<int64_t>(g.gTW.size());
    }

    v->feat = g.feat.empty() ? nullptr : g.feat.data();
    v->nfeat = static_cast<int64_t>(g.feat.size());

    if (grade) {
        *grade = g.grade.empty() ? nullptr : g.grade.data();
    }

    if (hbase) {
        *hbase = g.hbase;
    }
}

// seed_cpu into the handle; returns the node count (-1 on error)
int64_t ref_seed(void* h, const tn_relax_params* prm) {
    try {
        RefGrid* r = static_cast<RefGrid*>(h);
        r->seeds = tn::Nodes();
        tn::seed_cpu(r->g, to_params(prm), r->seeds);
        return static_cast<int64_t>(r->seeds.size());
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

void ref_seed_copy(void* h, float* P, uint16_t* lab, uint8_t* typ, uint16_t* part, uint16_t* part3) {
    const tn::Nodes& s = static_cast<RefGrid*>(h)->seeds;
    std::copy(s.P.begin(), s.P.end(), P);
    std::copy(s.lab.begin(), s.lab.end(), lab);
    std::copy(s.typ.begin(), s.typ.end(), typ);
    std::copy(s.part.begin(), s.part.end(), part);

    if (part3) {
        std::copy(s.part3.begin(), s.part3.end(), part3);
    }
}

// relax_cpu on the nodes (in place); the reference reads its rebuild divisor
// from V2M_REBUILD_DIV once per process, so only the default is accepted here
int ref_relax(void* h, const tn_relax_params* prm, tn_nodes_view* nd, tn_relax_stats* st) {
    try {
        if (prm->rebuild_div > 0 && prm->rebuild_div != 300) {
            throw std::runtime_error("ref_relax: only rebuild_div 300");
        }

        RefGrid* r = static_cast<RefGrid*>(h);
        const size_t n = static_cast<size_t>(nd->n);
        tn::Nodes x;
        x.P.assign(nd->P, nd->P + 3 * n);
        x.lab.assign(nd->lab, nd->lab + n);
        x.typ.assign(nd->typ, nd->typ + n);
        x.part.assign(nd->part, nd->part + 2 * n);
        x.part3.assign(n, V2M_NOLAB);
        tn::RelaxStats rs;
        tn::relax_cpu(r->g, to_params(prm), x, rs);
        std::copy(x.P.begin(), x.P.end(), nd->P);
        std::copy(x.typ.begin(), x.typ.end(), nd->typ);
        std::copy(x.part.begin(), x.part.end(), nd->part);
        st->iters = rs.iters;
        st->rebuilds = rs.rebuilds;
        st->last_move = rs.last_move;
        st->last_p99 = rs.last_p99;
        st->fire_resets = rs.fire_resets;
        st->fire_dt = rs.fire_dt;
        st->n_interior = static_cast<int64_t>(rs.n_interior);
        st->n_interface = static_cast<int64_t>(rs.n_interface);
        st->n_junction = static_cast<int64_t>(rs.n_junction);
        st->n_corner = static_cast<int64_t>(rs.n_corner);
        st->ms_seed = rs.ms_seed;
        st->ms_hash = rs.ms_hash;
        st->ms_force = rs.ms_force;
        st->ms_move = rs.ms_move;
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return 1;
    }
}

// the nodes as a tn::Nodes (part3 null: no third partners)
tn::Nodes nodes_of(int64_t n, const float* P, const uint16_t* lab, const uint8_t* typ, const uint16_t* part,
                   const uint16_t* part3) {
    tn::Nodes x;
    x.P.assign(P, P + 3 * n);
    x.lab.assign(lab, lab + n);
    x.typ.assign(typ, typ + n);
    x.part.assign(part, part + 2 * n);
    if (part3) {
        x.part3.assign(part3, part3 + n);
    } else {
        x.part3.assign(static_cast<size_t>(n), V2M_NOLAB);
    }
    return x;
}

void nodes_back(const tn::Nodes& x, float* P, uint16_t* lab, uint8_t* typ, uint16_t* part, uint16_t* part3) {
    std::copy(x.P.begin(), x.P.end(), P);
    std::copy(x.lab.begin(), x.lab.end(), lab);
    std::copy(x.typ.begin(), x.typ.end(), typ);
    std::copy(x.part.begin(), x.part.end(), part);
    if (part3) {
        std::copy(x.part3.begin(), x.part3.end(), part3);
    }
}

// thin_nodes (src/v2m_particles.cpp:1223) on the given nodes: the kept ones are
// written back, compacted in order; returns their count (-1 on error)
int64_t ref_thin_nodes(void* h, const tn_relax_params* prm, int64_t n, float* P, uint16_t* lab, uint8_t* typ,
                       uint16_t* part, uint16_t* part3) {
    try {
        RefGrid* r = static_cast<RefGrid*>(h);
        tn::Nodes x = nodes_of(n, P, lab, typ, part, part3);
        tn::thin_nodes(r->g, to_params(prm), x);
        nodes_back(x, P, lab, typ, part, part3);
        return static_cast<int64_t>(x.size());
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

// protect_features (src/v2m_particles.cpp:1002) with the grid's feature curves, as above
int64_t ref_protect_features(void* h, int64_t n, float* P, uint16_t* lab, uint8_t* typ, uint16_t* part,
                             uint16_t* part3) {
    try {
        RefGrid* r = static_cast<RefGrid*>(h);
        tn::Nodes x = nodes_of(n, P, lab, typ, part, part3);
        tn::protect_features(r->g, x);
        nodes_back(x, P, lab, typ, part, part3);
        return static_cast<int64_t>(x.size());
    } catch (const std::exception& e) {
        g_err = e.what();
        return -1;
    }
}

// the grid's feature curves (Grid::feat): their count, and the values when out is not null
int64_t ref_grid_feat(void* h, float* out) {
    const std::vector<float>& f = static_cast<RefGrid*>(h)->g.feat;
    if (out) {
        std::copy(f.begin(), f.end(), out);
    }
    return static_cast<int64_t>(f.size());
}

void ref_hash(void* h, const tn_relax_params* prm, tn_hash_info* hi) {
    int nkeys = 0;
    const refh::V2mHash H = hash_of(static_cast<RefGrid*>(h)->g, to_params(prm), &nkeys);
    std::memset(hi, 0, sizeof *hi);
    hi->b0 = H.b0;
    hi->ox = H.ox;
    hi->oy = H.oy;
    hi->oz = H.oz;
    hi->nlev = H.nlev;

    for (int L = 0; L < V2M_MAXLEV; ++L) {
        hi->off[L] = H.off[L];

        for (int a = 0; a < 3; ++a) {
            hi->dim[L][a] = H.dim[L][a];
        }
    }

    hi->nkeys = nkeys;
}

// The first rebuild of relax_cpu (src/v2m_particles.cpp:767-820) on the given
// nodes, through the reference's bodies: hn, keys, the stable counting sort,
// cell starts and the K-nearest lists.
void ref_rebuild(void* h, const tn_relax_params* prm, const tn_nodes_view* nd, int32_t* key, int32_t* sorted,
                 int32_t* cstart, int32_t* nbr, int32_t* nnb, float* hn) {
    const tn::Grid& g = static_cast<RefGrid*>(h)->g;
    const tn::RelaxParams p = to_params(prm);
    const refh::V2mDims d = dims_of(g);
    int nkeys = 0;
    const refh::V2mHash H = hash_of(g, p, &nkeys);
    const int n = static_cast<int>(nd->n);
    const float* P = nd->P;

    for (int i = 0; i < n; ++i) {
        hn[i] = refh::v2m_h_at(d, g.h.data(), P[3 * i], P[3 * i + 1], P[3 * i + 2]);
        key[i] = refh::v2m_bin_key(&H, refh::v2m_level_of(&H, (p.t + p.skin) * hn[i]), P[3 * i], P[3 * i + 1],
                                   P[3 * i + 2]);
    }

    std::fill(cstart, cstart + nkeys + 1, 0);

    for (int i = 0; i < n; ++i) {
        ++cstart[key[i] + 1];
    }

    for (int k = 0; k < nkeys; ++k) {
        cstart[k + 1] += cstart[k];
    }

    std::vector<int> cur(cstart, cstart + nkeys);

    for (int i = 0; i < n; ++i) {
        sorted[cur[key[i]]++] = i;
    }

    std::vector<float> Ps(static_cast<size_t>(n) * 4);

    for (int s = 0; s < n; ++s) {
        const int j = sorted[s];
        Ps[4 * s] = P[3 * j];
        Ps[4 * s + 1] = P[3 * j + 1];
        Ps[4 * s + 2] = P[3 * j + 2];
        Ps[4 * s + 3] = hn[j];
    }

    for (int i = 0; i < n; ++i) {
        float kds[V2M_K];
        int kids[V2M_K];
        refh::v2m_neighbors(&H, hn, P, nd->lab, nd->typ, cstart, sorted, Ps.data(), p.t, p.skin, i, nbr, nnb, kds, kids,
                            1);
    }
}

void ref_force(void* h, const tn_relax_params* prm, const tn_nodes_view* nd, const float* hn, const int32_t* nbr,
               const int32_t* nnb, float* F) {
    (void)h;
    const int n = static_cast<int>(nd->n);

    for (int i = 0; i < n; ++i) {
        refh::v2m_force(hn, nd->P, nd->typ, nbr, nnb, prm->fscale, prm->fsurf, i, &F[4 * i]);
    }
}

// One v2m_move per node (src/opencl/v2m_particle_body.cl:844-935) from the given
// forces and FIRE state; P, typ, part, hn, V updated in place, pw and mv written.
void ref_move(void* h, const tn_relax_params* prm, tn_nodes_view* nd, float* hn, const float* F, float* V, float fdt,
              float falpha, float* pw, float* mv) {
    const RefGrid* r = static_cast<RefGrid*>(h);
    const tn::Grid& g = r->g;
    const refh::V2mDims d = dims_of(g);
    const int n = static_cast<int>(nd->n);
    const int fire = prm->fire && !prm->voxel_trap ? 1 : 0;

    for (int i = 0; i < n; ++i) {
        mv[i] = refh::v2m_move(d, g.L->data(), g.bl_cnt.data(), g.bl_lab.data(), g.bl_slot.data(), g.phi.data(), g.gI,
                               g.gTW.data(), g.gm, g.h.data(), F, prm->dt, prm->maxstep, prm->snap,
                               prm->voxel_trap ? 1 : 0, i, nd->P, nd->lab, nd->typ, nd->part, hn, fire, V, fdt, falpha,
                               pw);
    }
}

// Seeds from arbitrary points: each labelled by v2m_label_of (voxel mode off), the
// exterior ones (label 0) dropped, the rest classified by v2m_seed_classify
// (src/opencl/v2m_seed_body.cl:116-175, 511-599). Writes the kept nodes in input
// order (P, lab, typ, part) and returns their count.
int64_t ref_classify_points(void* h, int64_t m, const float* X, float* P, uint16_t* lab, uint8_t* typ,
                            uint16_t* part) {
    const RefGrid* r = static_cast<RefGrid*>(h);
    const tn::Grid& g = r->g;
    const refh::V2mDims d = dims_of(g);
    int64_t n = 0;

    for (int64_t k = 0; k < m; ++k) {
        const float* x = X + 3 * k;
        int sec;
        float mg;
        const int l = refh::v2m_label_of(d, g.L->data(), g.bl_cnt.data(), g.bl_lab.data(), g.bl_slot.data(),
                                         g.phi.data(), g.gI, g.gTW.data(), g.gm, 0, x[0], x[1], x[2], &sec, &mg);

        if (l == 0 || l == V2M_NOLAB) {
            continue;
        }

        P[3 * n] = x[0];
        P[3 * n + 1] = x[1];
        P[3 * n + 2] = x[2];
        lab[n] = static_cast<uint16_t>(l);
        ++n;
    }

    for (int64_t i = 0; i < n; ++i) {
        refh::v2m_seed_classify(d, g.L->data(), g.bl_cnt.data(), g.bl_lab.data(), g.bl_slot.data(), g.phi.data(),
                                g.gI, g.gTW.data(), g.gm, g.h.data(), static_cast<int>(i), P, lab, typ, part);
    }

    return n;
}

// point queries on the field, for unit fixtures: phi_l and psi_ab (+ gradient)
void ref_field_points(void* h, int a, int b, int64_t m, const float* X, float* phi_a, float* psi, float* grad,
                      int32_t* label, int32_t* second) {
    const RefGrid* r = static_cast<RefGrid*>(h);
    const tn::Grid& g = r->g;
    const refh::V2mDims d = dims_of(g);

    for (int64_t k = 0; k < m; ++k) {
        const float* x = X + 3 * k;
        phi_a[k] = refh::v2m_phi_f(d, g.L->data(), g.bl_cnt.data(), g.bl_lab.data(), g.bl_slot.data(), g.phi.data(),
                                   g.gI, g.gTW.data(), g.gm, a, x[0], x[1], x[2]);
        psi[k] = refh::v2m_psi_grad(d, g.L->data(), g.bl_cnt.data(), g.bl_lab.data(), g.bl_slot.data(), g.phi.data(),
                                    g.gI, g.gTW.data(), g.gm, a, b, x[0], x[1], x[2], grad + 3 * k);
        int sec;
        float mg;
        label[k] = refh::v2m_label_of(d, g.L->data(), g.bl_cnt.data(), g.bl_lab.data(), g.bl_slot.data(), g.phi.data(),
                                      g.gI, g.gTW.data(), g.gm, 0, x[0], x[1], x[2], &sec, &mg);
        second[k] = sec;
    }
}

// tet_quality (src/v2m_tetra.cpp:100-150) of the tet p[12] (its 4 corners): out[3] =
// (the minimum dihedral in degrees, the Joe-Liu quality, the volume)
void ref_tet_quality(const double* p, double* out) {
    const double* q[4] = { p, p + 3, p + 6, p + 9 };
    tn::tet_quality(q, out[0], out[1], out[2]);
}

// TetMesh::tetrahedrize (third_party/cdt/src/delaunay.cpp) of n points X[3n], in the
// canonical form of tn::canonicalize_tets (src/v2m_tetra.cpp:3928): the finite tets'
// corners, 4 each, into tets when cap >= their count. Returns the finite tet count.
int64_t ref_delaunay(const double* X, int64_t n, int32_t* tets, int64_t cap) {
    try {
        ::TetMesh tm;
        tm.init_vertices(X, static_cast<uint32_t>(n));
        tm.tetrahedrize();
        tn::canonicalize_tets(tm);
        int64_t m = 0;
        for (uint64_t t = 0; t < tm.numTets(); ++t) {
            const uint32_t* v = &tm.tet_node[4 * t];
            if (v[3] == INFINITE_VERTEX) {
                continue;
            }
            if (tets && m < cap) {
                for (int k = 0; k < 4; ++k) {
                    tets[4 * m + k] = static_cast<int32_t>(v[k]);
                }
            }
            ++m;
        }
        return m;
    } catch (...) {
        return -1;
    }
}

}  // extern "C"
