// v2m_hip.cpp -- v2mesh's node-placement entry points on an MI355X (see v2m_hip.h).
//
// Compiled inside v2mesh (its src/ headers on the include path) with
// include/trussnet_amd.hpp, linked with -ltrussnet_amd. The views point at the
// reference's own vectors: no host copies of the grid or the nodes.
#include "v2m_hip.h"

#include <algorithm>
#include <cstdlib>
#include <stdexcept>
#include <string>

#include "trussnet_amd.hpp"

namespace tn {

namespace {

tn_grid_view grid_view(const Grid& g) {
    if (!g.L) {
        throw std::runtime_error("relax_hip: Grid without labels");
    }
    tn_grid_view v{};
    v.nx = g.nx;
    v.ny = g.ny;
    v.nz = g.nz;
    v.nbx = g.nbx;
    v.nby = g.nby;
    v.nbz = g.nbz;
    for (int a = 0; a < 3; ++a) {
        v.vs[a] = g.vs[a];
    }
    v.L = g.L->data();
    v.bl_cnt = g.bl_cnt.data();
    v.bl_lab = g.bl_lab.data();
    v.bl_slot = g.bl_slot.data();
    v.phi = g.phi.data();
    v.nslot = (int64_t)g.slot_label.size();
    v.h = g.h.data();
    v.hmin = g.hmin;
    v.hmax = g.hmax;
    v.gI = g.gI;
    v.gTW = g.gTW.empty() ? nullptr : g.gTW.data();
    v.ngTW = (int64_t)g.gTW.size();
    v.gm = g.gm;
    v.feat = g.feat.empty() ? nullptr : g.feat.data();
    v.nfeat = (int64_t)g.feat.size();
    return v;
}

tn_relax_params params(const RelaxParams& p) {
    tn_relax_params c = tnamd::default_params();
    c.nseed = p.nseed;
    c.t = p.t;
    c.skin = p.skin;
    c.fscale = p.fscale;
    c.fsurf = p.fsurf;
    c.dt = p.dt;
    c.maxstep = p.maxstep;
    c.snap = p.snap;
    c.max_iters = p.max_iters;
    c.dptol = p.dptol;
    c.jseed = p.jseed;
    c.corners = p.corners ? 1 : 0;
    c.voxel_trap = p.voxel_trap ? 1 : 0;
    c.thin = p.thin;
    c.fire = p.fire ? 1 : 0;
    c.fire_dtmax = p.fire_dtmax;
    c.verbose = p.verbose ? 1 : 0;
    const char* div = std::getenv("V2M_REBUILD_DIV");   // as relax_cpu reads it (src/v2m_particles.cpp:731)
    c.rebuild_div = div ? std::atoi(div) : 300;
    return c;
}

tn_nodes_view nodes_view(Nodes& nd) {
    const size_t n = nd.size();
    if (nd.P.size() != 3 * n || nd.typ.size() != n || nd.part.size() != 2 * n) {
        throw std::runtime_error("relax_hip: inconsistent Nodes arrays");
    }
    return tn_nodes_view{ (int64_t)n, nd.P.data(), nd.lab.data(), nd.typ.data(), nd.part.data() };
}

// A call that compacts the node arrays in place (thin_nodes, protect_features), then
// the vectors resized to the nodes kept, as the reference's own functions leave them.
// Returns the count removed.
template <class Call>
size_t compacting(Nodes& nd, const char* what, Call call) {
    const size_t n = nd.size();
    if (nd.P.size() != 3 * n || nd.typ.size() != n || nd.part.size() != 2 * n ||
        (!nd.part3.empty() && nd.part3.size() != n)) {
        throw std::runtime_error(std::string(what) + ": inconsistent Nodes arrays");
    }
    int64_t kept = (int64_t)n;
    tnamd::check(call((int64_t)n, nd.part3.empty() ? nullptr : nd.part3.data(), &kept), what);
    const size_t m = (size_t)kept;
    nd.P.resize(3 * m);
    nd.lab.resize(m);
    nd.typ.resize(m);
    nd.part.resize(2 * m);
    if (!nd.part3.empty()) {
        nd.part3.resize(m);
    }
    return n - m;
}

void stats_from(const tn_relax_stats& s, RelaxStats& st) {
    st.iters = s.iters;
    st.rebuilds = s.rebuilds;
    st.last_move = s.last_move;
    st.last_p99 = s.last_p99;
    st.fire_resets = s.fire_resets;
    st.fire_dt = s.fire_dt;
    st.n_interior = (size_t)s.n_interior;
    st.n_interface = (size_t)s.n_interface;
    st.n_junction = (size_t)s.n_junction;
    st.n_corner = (size_t)s.n_corner;
    st.ms_hash = s.ms_hash;
    st.ms_force = s.ms_force;
    st.ms_move = s.ms_move;
}

}  // namespace

void relax_hip(const Grid& g, const RelaxParams& prm, Nodes& nd, RelaxStats& st, int device) {
    const tn_grid_view gv = grid_view(g);
    const tn_relax_params pc = params(prm);
    tn_nodes_view nv = nodes_view(nd);
    tn_relax_stats s{};
    tnamd::relax(gv, pc, nv, s, device);   // throws tnamd::Error, a std::runtime_error
    stats_from(s, st);
}

void relax_hip_batch(const std::vector<const Grid*>& g, const RelaxParams& prm, std::vector<Nodes*>& nd,
                     std::vector<RelaxStats>& st, int device) {
    if (g.size() != nd.size()) {
        throw std::runtime_error("relax_hip_batch: one Nodes per Grid");
    }
    const size_t m = g.size();
    std::vector<tn_grid_view> gv(m);
    std::vector<tn_relax_params> pc(m, params(prm));
    std::vector<tn_nodes_view> nv(m);
    std::vector<tn_relax_stats> s(m);
    for (size_t k = 0; k < m; ++k) {
        gv[k] = grid_view(*g[k]);
        nv[k] = nodes_view(*nd[k]);
    }
    tnamd::check(tn_relax_batch(gv.data(), pc.data(), nv.data(), s.data(), (int32_t)m, device), "tn_relax_batch");
    st.assign(m, RelaxStats());
    for (size_t k = 0; k < m; ++k) {
        stats_from(s[k], st[k]);
    }
}

void seed_hip(const Grid& g, const RelaxParams& prm, Nodes& nd, int device) {
    const tn_grid_view gv = grid_view(g);
    const tn_relax_params pc = params(prm);
    int64_t n = 0;
    tnamd::check(tn_seed_count(&gv, g.grade.data(), &pc, device, &n), "tn_seed_count");
    Nodes out;
    out.P.resize(3 * (size_t)n);
    out.lab.resize(n);
    out.typ.resize(n);
    out.part.resize(2 * (size_t)n);
    out.part3.resize(n);
    tnamd::check(tn_seed(&gv, g.grade.data(), &pc, device, n, out.P.data(), out.lab.data(), out.typ.data(),
                         out.part.data(), out.part3.data()),
                 "tn_seed");
    nd = std::move(out);
}

size_t thin_hip(const Grid& g, const RelaxParams& prm, Nodes& nd, int device) {
    const tn_grid_view gv = grid_view(g);
    const tn_relax_params pc = params(prm);
    return compacting(nd, "tn_thin_nodes", [&](int64_t n, uint16_t* part3, int64_t* kept) {
        return tn_thin_nodes(&gv, &pc, n, nd.P.data(), nd.lab.data(), nd.typ.data(), nd.part.data(), part3, kept,
                             device);
    });
}

size_t protect_hip(const Grid& g, Nodes& nd, int device) {
    const tn_grid_view gv = grid_view(g);
    return compacting(nd, "tn_protect_features", [&](int64_t n, uint16_t* part3, int64_t* kept) {
        return tn_protect_features(&gv, g.feat.empty() ? nullptr : g.feat.data(), (int64_t)g.feat.size(), n,
                                   nd.P.data(), nd.lab.data(), nd.typ.data(), nd.part.data(), part3, kept, device);
    });
}

void build_grid_hip(const LabelVolume& lv, const GridParams& p, Grid& g, int device) {
    tn_volume_view v{};
    v.nx = lv.nx;
    v.ny = lv.ny;
    v.nz = lv.nz;
    for (int a = 0; a < 3; ++a) {
        v.vs[a] = (float)lv.voxelsize[a];
    }
    v.L = lv.data.data();
    const bool tpm = !lv.prob.empty() && lv.nprob > 0;
    if (tpm) {
        v.prob = lv.prob.data();
        v.nprob = lv.nprob;
    } else if (!lv.gray.empty() && !lv.thresholds.empty()) {
        v.gray = lv.gray.data();
        v.thresholds = lv.thresholds.data();
        v.nthr = (int32_t)lv.thresholds.size();
    }
    if (!lv.sdf.empty()) {
        throw tnamd::Error(TN_ERR_UNSUPPORTED, "build_grid_hip: shape inputs are built on the host");
    }
    if (!p.hvox.empty()) {
        v.hvox = p.hvox.data();
    }
    tn_grid_params c = tnamd::default_grid_params();
    c.sigma = p.sigma;
    c.sigma_thin = p.sigma_thin;
    c.thin_lo = p.thin_lo;
    c.thin_hi = p.thin_hi;
    c.thick = p.thick;
    c.thin_floor = p.thin_floor;
    c.preserve = p.preserve;
    c.sigma_curv = p.sigma_curv;
    c.hbase = p.hbase;
    c.hmin = p.hmin;
    c.hmax = p.hmax;
    c.K = p.K;
    c.g = p.g;
    c.hlab = p.hlab.empty() ? nullptr : p.hlab.data();
    c.nhlab = (int32_t)p.hlab.size();
    c.iglob = p.isize.glob;
    c.ilab = p.isize.lab.empty() ? nullptr : p.isize.lab.data();
    c.nilab = (int32_t)p.isize.lab.size();
    c.ipair = p.isize.pair.empty() ? nullptr : p.isize.pair.data();
    c.npair = (int32_t)(p.isize.pair.size() / 3);
    const tnamd::DeviceGrid dg(v, c, device);   // released on every path out
    const tn_grid_info in = dg.info();
    Grid o;
    o.nx = in.nx;
    o.ny = in.ny;
    o.nz = in.nz;
    o.nbx = in.nbx;
    o.nby = in.nby;
    o.nbz = in.nbz;
    for (int a = 0; a < 3; ++a) {
        o.vs[a] = v.vs[a];
    }
    o.L = &lv.data;
    o.R = in.R;
    o.sigma = p.sigma;
    o.nlab = lv.maxlabel + 1;
    o.hbase = in.hbase;
    o.hmin = in.hmin;
    o.hmax = in.hmax;
    o.hcurv = in.hcurv;
    o.prob_fields = in.prob_fields != 0;
    o.limit_sweeps = in.limit_sweeps;
    o.overflow_bricks = in.overflow_bricks;
    const size_t nb = (size_t)in.nbx * in.nby * in.nbz, nv = (size_t)in.nx * in.ny * in.nz;
    o.bl_cnt.resize(nb);
    o.bl_lab.resize((size_t)TN_BL * nb);
    o.bl_slot.resize(nb);
    o.phi.resize((size_t)in.nslot * TN_SLOT);
    o.h.resize(nv);
    o.grade.resize(nv);
    o.gm = in.gm;
    o.gTW.assign(in.gm ? 2 * (size_t)in.gm : 1, 0.0f);
    dg.download(o.bl_cnt.data(), o.bl_lab.data(), o.bl_slot.data(), o.phi.data(), o.h.data(), o.grade.data(),
                in.gm ? o.gTW.data() : nullptr);
    if (in.gm) {
        o.gI = lv.gray.data();
        o.gray_w.assign(o.gTW.begin() + in.gm, o.gTW.end());
    }
    // the slots' (brick, label): the mixed bricks' labels in brick order
    for (size_t b = 0; b < nb; ++b) {
        if (o.bl_slot[b] >= 0) {
            for (int i = 0; i < std::min(o.bl_cnt[b], TN_BL); ++i) {
                o.slot_brick.push_back((int)b);
                o.slot_label.push_back(o.bl_lab[(size_t)TN_BL * b + i]);
            }
        }
    }
    g = std::move(o);
}

}  // namespace tn
