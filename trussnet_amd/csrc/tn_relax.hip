// tn_relax.hip -- host driver of the gfx950 relaxation and its C ABI
// (include/trussnet_amd.h).
//
// The loop is relax_cl's (src/v2m_gpu.cpp:116-553) and relax_cpu's
// (src/v2m_particles.cpp:728-976): rebuild, then per iteration force -> move ->
// stats -> (own-list refreshes) -> control, rebuilding when the control kernel
// asks. Unlike the OpenCL host, FIRE's update, the p99 and both decisions run on
// the device; the host reads back one small control block per iteration.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "tn_kernels.hpp"

namespace {

thread_local std::string g_err;

struct TnError : std::runtime_error {
    int code;
    TnError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define TN_HIP(call)                                                                                   \
    do {                                                                                               \
        const hipError_t e_ = (call);                                                                  \
        if (e_ != hipSuccess) {                                                                        \
            throw TnError(TN_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_));              \
        }                                                                                              \
    } while (0)

inline unsigned blocks_for(int64_t n, int b) {
    return (unsigned)((n + b - 1) / b);
}

template <class F>
int guarded(F&& f) {
    try {
        f();
        return TN_OK;
    } catch (const TnError& e) {
        g_err = e.what();
        return e.code;
    } catch (const std::bad_alloc&) {
        g_err = "out of host memory";
        return TN_ERR_INTERNAL;
    } catch (const std::exception& e) {
        g_err = e.what();
        return TN_ERR_INTERNAL;
    }
}

void arg_check(bool ok, const std::string& what) {
    if (!ok) {
        throw TnError(TN_ERR_ARG, what);
    }
}

// the hash geometry of relax_cpu (src/v2m_particles.cpp:738-761)
tnd::Hash make_hash(const tn_grid_view* g, const tn_relax_params* prm, int64_t* nkeys_out) {
    tnd::Hash H;
    std::memset(&H, 0, sizeof H);
    H.b0 = (prm->t + prm->skin) * g->hmin;
    H.ox = -0.5f * g->vs[0];
    H.oy = -0.5f * g->vs[1];
    H.oz = -0.5f * g->vs[2];
    const float ext[3] = { (float)g->nx * g->vs[0], (float)g->ny * g->vs[1], (float)g->nz * g->vs[2] };
    H.nlev = 1;
    while (H.nlev < TN_MAXLEV && H.b0 * (float)(1 << (H.nlev - 1)) < (prm->t + prm->skin) * g->hmax) {
        ++H.nlev;
    }
    int64_t nkeys = 0;
    for (int L = 0; L < H.nlev; ++L) {
        H.off[L] = (int)nkeys;
        const float bL = H.b0 * (float)(1 << L);
        for (int a = 0; a < 3; ++a) {
            H.dim[L][a] = std::max(1, (int)std::ceil(ext[a] / bL));
        }
        nkeys += (int64_t)H.dim[L][0] * H.dim[L][1] * H.dim[L][2];
        arg_check(nkeys < (int64_t)1 << 30, "hash: too many cells (hmin too small for the grid)");
    }
    H.nkeys = (int)nkeys;
    *nkeys_out = nkeys;
    return H;
}

// thr[e + 21]: the smallest float m below 2^e with floor(log2f(m)) == e (the host's
// log2f, as the CPU path's std::log2 on a float), or 2^e if there is none
void hist_thresholds(float* thr) {
    for (int e = -21; e <= 13; ++e) {
        const float p2 = std::ldexp(1.0f, e);
        float m = std::nextafter(p2, 0.0f);
        if (std::floor(std::log2(m)) < (float)e) {
            thr[e + 21] = p2;
            continue;
        }
        for (;;) {
            const float lo = std::nextafter(m, 0.0f);
            if (std::floor(std::log2(lo)) < (float)e) {
                break;
            }
            m = lo;
        }
        thr[e + 21] = m;
    }
}

__global__ void k_init_nodes(tnd::Field f, const float* __restrict__ P, int n, float4* __restrict__ X) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const float x = P[3 * i], y = P[3 * i + 1], z = P[3 * i + 2];
        X[i] = make_float4(x, y, z, tnd::size_at(f, x, y, z));
    }
}

__global__ void k_iota(int32_t* __restrict__ a, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        a[i] = i;
    }
}

}  // namespace

struct tn_plan {
    int device = 0;
    hipStream_t stream = nullptr;
    tn_relax_params prm{};
    int n = 0, fire = 0, rebuild_div = 300;
    int64_t nkeys = 0;
    tnd::Field f{};
    tnd::Hash H{};
    float tskin = 0.0f;
    bool profile = false;
    // device buffers
    std::vector<void*> owned;
    float4 *dX0 = nullptr, *dX = nullptr, *dV = nullptr, *dF = nullptr, *dP0 = nullptr, *dPs = nullptr;
    uint8_t *dTyp0 = nullptr, *dTyp = nullptr;
    uint32_t *dPart0 = nullptr, *dPart = nullptr;
    uint16_t* dLab = nullptr;
    float *dMv = nullptr, *dPw = nullptr, *dThr = nullptr;
    int32_t *dKey = nullptr, *dCstart = nullptr, *dSorted = nullptr, *dNbrT = nullptr, *dNnb = nullptr;
    int32_t *dOrder = nullptr, *dQueue = nullptr, *dBhist = nullptr, *dBsum = nullptr;
    int32_t* dRk[2] = { nullptr, nullptr };
    int32_t* dRv[2] = { nullptr, nullptr };
    double2* dPartials = nullptr;
    tnk::Ctl* dCtl = nullptr;
    tnk::Ctl* hCtl = nullptr;   // pinned
    int nparts = 0, radix_passes = 0, radix_nblk = 0;
    // host-side loop state
    int rebuilds = 0;
    bool finished = false;
    double ms_hash = 0, ms_force = 0, ms_move = 0;
    hipEvent_t ev[4] = { nullptr, nullptr, nullptr, nullptr };

    template <class T>
    T* alloc(size_t count) {
        void* p = nullptr;
        TN_HIP(hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T)));
        owned.push_back(p);
        return static_cast<T*>(p);
    }

    ~tn_plan() {
        if (device >= 0) {
            hipSetDevice(device);
        }
        for (void* p : owned) {
            hipFree(p);
        }
        if (hCtl) {
            hipHostFree(hCtl);
        }
        for (hipEvent_t& e : ev) {
            if (e) {
                hipEventDestroy(e);
            }
        }
        if (stream) {
            hipStreamDestroy(stream);
        }
    }

    void scan(const int32_t* in, int32_t* out, int64_t m) {
        const unsigned nb = blocks_for(m, tnk::kScanTile);
        tnk::k_scan_reduce<<<nb, tnk::kBlock, 0, stream>>>(in, m, dBsum);
        tnk::k_scan_top<<<1, tnk::kBlock, 0, stream>>>(dBsum, nb);
        tnk::k_scan_down<<<nb, tnk::kBlock, 0, stream>>>(in, out, m, dBsum);
    }

    tnd::TrussIn truss_in() const {
        tnd::TrussIn in;
        in.X = dX;
        in.lab = dLab;
        in.typ = dTyp;
        in.cstart = dCstart;
        in.sorted = dSorted;
        in.Ps = dPs;
        in.t = prm.t;
        in.skin = prm.skin;
        return in;
    }

    // the rebuild lambda (src/v2m_particles.cpp:771-820) and the type order
    // (src/v2m_gpu.cpp:288-302)
    void rebuild() {
        if (profile) {
            TN_HIP(hipEventRecord(ev[0], stream));
        }
        const unsigned nb = blocks_for(n, tnk::kBlock);
        tnk::k_fill_i32<<<blocks_for(nkeys + 1, tnk::kBlock), tnk::kBlock, 0, stream>>>(dCstart, nkeys + 1, 0);
        tnk::k_keys<<<nb, tnk::kBlock, 0, stream>>>(f, H, dX, n, tskin, dKey, dCstart);
        scan(dCstart, dCstart, nkeys + 1);
        if (radix_passes == 0) {
            k_iota<<<nb, tnk::kBlock, 0, stream>>>(dSorted, n);
        }
        for (int p = 0; p < radix_passes; ++p) {
            const int32_t* kin = p == 0 ? dKey : dRk[(p - 1) & 1];
            const int32_t* vin = p == 0 ? nullptr : dRv[(p - 1) & 1];
            int32_t* kout = dRk[p & 1];
            int32_t* vout = p == radix_passes - 1 ? dSorted : dRv[p & 1];
            tnk::k_radix_hist<<<radix_nblk, tnk::kBlock, 0, stream>>>(kin, n, 8 * p, radix_nblk, dBhist);
            scan(dBhist, dBhist, (int64_t)256 * radix_nblk);
            tnk::k_radix_scatter<<<radix_nblk, tnk::kBlock, 0, stream>>>(kin, vin, n, 8 * p, radix_nblk, dBhist, kout,
                                                                          vout);
        }
        tnk::k_gather<<<nb, tnk::kBlock, 0, stream>>>(dSorted, dX, n, dPs, dP0);
        tnk::k_neighbors<<<blocks_for(n, tnk::kNbrBlock), tnk::kNbrBlock, 0, stream>>>(H, truss_in(), n, dNbrT, dNnb);
        tnk::k_order_begin<<<1, 64, 0, stream>>>(dCtl);
        tnk::k_order_count<<<nb, tnk::kBlock, 0, stream>>>(dTyp, n, dCtl);
        tnk::k_order_scatter<<<nb, tnk::kBlock, 0, stream>>>(dTyp, n, dCtl, dOrder);
        ++rebuilds;
        if (profile) {
            TN_HIP(hipEventRecord(ev[1], stream));
            TN_HIP(hipEventSynchronize(ev[1]));
            float ms = 0;
            TN_HIP(hipEventElapsedTime(&ms, ev[0], ev[1]));
            ms_hash += ms;
        }
    }

    // device state back to the seeds, then the first rebuild
    void reset() {
        TN_HIP(hipSetDevice(device));
        TN_HIP(hipMemcpyAsync(dX, dX0, sizeof(float4) * n, hipMemcpyDeviceToDevice, stream));
        TN_HIP(hipMemcpyAsync(dTyp, dTyp0, n, hipMemcpyDeviceToDevice, stream));
        TN_HIP(hipMemcpyAsync(dPart, dPart0, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, stream));
        tnk::Ctl c;
        std::memset(&c, 0, sizeof c);
        c.dt0 = std::sqrt(prm.dt);   // FireCtl's constructor (src/v2m_particles.h:50-51)
        c.dt = c.dt0;
        c.dtmax = prm.fire_dtmax * c.dt0;
        c.alpha = 0.1f;
        c.zero_v = 1;   // the velocities start at rest
        *hCtl = c;
        TN_HIP(hipMemcpyAsync(dCtl, hCtl, sizeof c, hipMemcpyHostToDevice, stream));
        rebuilds = 0;
        finished = false;
        ms_hash = ms_force = ms_move = 0;
        if (n > 0) {
            rebuild();
        }
        TN_HIP(hipStreamSynchronize(stream));
    }

    // one iteration (src/v2m_particles.cpp:828-946); returns false once converged
    bool iterate() {
        const unsigned nb = blocks_for(n, tnk::kBlock);
        if (profile) {
            TN_HIP(hipEventRecord(ev[0], stream));
        }
        tnk::k_iter_begin<<<1, 64, 0, stream>>>(dCtl);
        tnk::k_force<<<nb, tnk::kBlock, 0, stream>>>(dX, dTyp, dNbrT, dNnb, prm.fscale, prm.fsurf, n, dF);
        if (profile) {
            TN_HIP(hipEventRecord(ev[1], stream));
        }
        tnk::MoveArgs ma{ prm.dt, prm.maxstep, prm.snap, fire };
        tnk::k_move<<<nb, tnk::kBlock, 0, stream>>>(f, ma, dCtl, dOrder, n, dF, dX, dV, dTyp, dPart, dLab, dMv, dPw);
        tnk::k_stats<<<nb, tnk::kBlock, 0, stream>>>(dX, dP0, dMv, dPw, n, prm.skin, fire, dThr, dCtl, dQueue,
                                                     dPartials);
        tnk::k_refresh<<<256, tnk::kNbrBlock, 0, stream>>>(H, truss_in(), n, dQueue, dCtl, dNbrT, dNnb, dP0);
        tnk::CtlArgs ca{ n, prm.max_iters, fire, rebuild_div, prm.dptol, nparts };
        tnk::k_control<<<1, tnk::kBlock, 0, stream>>>(ca, dCtl, dPartials, dPw);
        TN_HIP(hipMemcpyAsync(hCtl, dCtl, sizeof(tnk::Ctl), hipMemcpyDeviceToHost, stream));
        if (profile) {
            TN_HIP(hipEventRecord(ev[2], stream));
        }
        TN_HIP(hipStreamSynchronize(stream));
        if (profile) {
            float a = 0, b = 0;
            TN_HIP(hipEventElapsedTime(&a, ev[0], ev[1]));
            TN_HIP(hipEventElapsedTime(&b, ev[1], ev[2]));
            ms_force += a;
            ms_move += b;
        }
        if (hCtl->done) {
            finished = true;
            return false;
        }
        if (hCtl->rebuild) {
            rebuild();
        }
        return true;
    }

    void run_iters(int iters) {
        for (int k = 0; k < iters && !finished && hCtl->it < prm.max_iters; ++k) {
            if (!iterate()) {
                break;
            }
        }
        TN_HIP(hipStreamSynchronize(stream));
    }

    void fill_stats(tn_relax_stats* st) {
        if (!st) {
            return;
        }
        std::memset(st, 0, sizeof *st);
        const tnk::Ctl& c = *hCtl;
        st->iters = c.it;
        st->rebuilds = rebuilds;
        st->last_move = c.last_move;
        st->last_p99 = c.last_p99;
        st->fire_resets = c.fire_resets;
        st->fire_dt = c.fire_dt;
        st->ms_hash = ms_hash;
        st->ms_force = ms_force;
        st->ms_move = ms_move;
        if (n > 0) {
            std::vector<uint8_t> typ(n);
            TN_HIP(hipMemcpy(typ.data(), dTyp, n, hipMemcpyDeviceToHost));
            for (uint8_t t : typ) {
                st->n_interior += t == TN_INTERIOR;
                st->n_interface += t == TN_INTERFACE;
                st->n_junction += t == TN_JUNCTION;
                st->n_corner += t == TN_CORNER;
            }
        }
    }
};

namespace {

int pick_device(int32_t device) {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) {
        throw TnError(TN_ERR_NODEVICE, "no HIP device");
    }
    const int d = device < 0 ? 0 : device;
    if (d >= count) {
        throw TnError(TN_ERR_NODEVICE, "HIP device " + std::to_string(device) + " out of range (" +
                                           std::to_string(count) + " devices)");
    }
    return d;
}

void check_grid(const tn_grid_view* g) {
    arg_check(g != nullptr, "null grid");
    if (g->gm != 0) {
        throw TnError(TN_ERR_UNSUPPORTED, "gray-scale and shape fields (gm != 0) are not run on the device");
    }
    arg_check(g->nx > 0 && g->ny > 0 && g->nz > 0, "grid: empty");
    arg_check(g->nbx == (g->nx + 7) / 8 && g->nby == (g->ny + 7) / 8 && g->nbz == (g->nz + 7) / 8,
              "grid: brick counts must be ceil(n / 8) per axis");
    arg_check(g->vs[0] > 0 && g->vs[1] > 0 && g->vs[2] > 0, "grid: voxel sizes must be positive");
    arg_check(g->hmin > 0 && g->hmax >= g->hmin, "grid: need 0 < hmin <= hmax");
    arg_check(g->L && g->bl_cnt && g->bl_lab && g->bl_slot && g->h, "grid: null array");
    arg_check(g->nslot >= 0 && (g->nslot == 0 || g->phi), "grid: null phi");
    const int64_t nb = (int64_t)g->nbx * g->nby * g->nbz;
    for (int64_t b = 0; b < nb; ++b) {   // every slot a lookup can reach lies in phi
        const int cnt = g->bl_cnt[b], s0 = g->bl_slot[b];
        arg_check(cnt >= 0 && cnt <= TN_BL + 1, "grid: bl_cnt out of range");
        if (s0 >= 0) {
            arg_check((int64_t)s0 + std::min(cnt, TN_BL) <= g->nslot, "grid: bl_slot beyond phi");
        }
    }
}

void check_params(const tn_relax_params* prm) {
    arg_check(prm != nullptr, "null params");
    if (prm->voxel_trap) {
        throw TnError(TN_ERR_UNSUPPORTED, "the voxel trap mode is not run on the device");
    }
    arg_check(prm->t + prm->skin > 0, "params: t + skin must be positive");
    arg_check(prm->max_iters >= 0, "params: max_iters must be >= 0");
}

void check_nodes(const tn_nodes_view* nd) {
    arg_check(nd != nullptr, "null nodes");
    arg_check(nd->n >= 0 && nd->n < ((int64_t)1 << 31) / TN_K, "nodes: count out of range");
    if (nd->n > 0) {
        arg_check(nd->P && nd->lab && nd->typ && nd->part, "nodes: null array");
    }
}

}  // namespace

extern "C" {

void tn_relax_params_default(tn_relax_params* p) {
    if (!p) {
        return;
    }
    std::memset(p, 0, sizeof *p);
    p->nseed = 8;
    p->t = 1.3f;
    p->skin = 0.3f;
    p->fscale = 1.2f;
    p->fsurf = 1.0f;
    p->dt = 0.5f;
    p->maxstep = 0.2f;
    p->snap = 0.5f;
    p->max_iters = 500;
    p->dptol = 2e-3f;
    p->jseed = 0.8f;
    p->corners = 1;
    p->voxel_trap = 0;
    p->thin = 0.0f;
    p->fire = 1;
    p->fire_dtmax = 2.0f;
    p->verbose = 0;
    p->rebuild_div = 300;
}

const char* tn_last_error(void) {
    return g_err.c_str();
}

int tn_device_count(int32_t* count) {
    return guarded([&] {
        arg_check(count != nullptr, "null count");
        int c = 0;
        if (hipGetDeviceCount(&c) != hipSuccess) {
            c = 0;
        }
        *count = c;
        if (c <= 0) {
            throw TnError(TN_ERR_NODEVICE, "no HIP device");
        }
    });
}

int tn_plan_create(const tn_grid_view* g, const tn_relax_params* prm, const tn_nodes_view* seeds, int32_t device,
                   tn_plan** out) {
    return guarded([&] {
        arg_check(out != nullptr, "null plan pointer");
        *out = nullptr;
        check_grid(g);
        check_params(prm);
        check_nodes(seeds);
        const int dev = pick_device(device);
        TN_HIP(hipSetDevice(dev));
        std::unique_ptr<tn_plan> p(new tn_plan);
        p->device = dev;
        p->prm = *prm;
        p->n = (int)seeds->n;
        p->fire = prm->fire && !prm->voxel_trap ? 1 : 0;
        p->rebuild_div = prm->rebuild_div > 0 ? prm->rebuild_div : 300;
        p->tskin = prm->t + prm->skin;
        p->profile = prm->verbose || std::getenv("TN_PROFILE") != nullptr;
        p->H = make_hash(g, prm, &p->nkeys);
        TN_HIP(hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking));
        for (hipEvent_t& e : p->ev) {
            TN_HIP(hipEventCreate(&e));
        }
        const int n = p->n;
        const size_t nv = (size_t)g->nx * g->ny * g->nz;
        const size_t nb = (size_t)g->nbx * g->nby * g->nbz;
        // the grid
        uint16_t* dL = p->alloc<uint16_t>(nv);
        int32_t* dCnt = p->alloc<int32_t>(nb);
        uint16_t* dBl = p->alloc<uint16_t>(nb * TN_BL);
        int32_t* dSlot = p->alloc<int32_t>(nb);
        float* dPhi = p->alloc<float>((size_t)g->nslot * TN_SLOT);
        float* dH = p->alloc<float>(nv);
        TN_HIP(hipMemcpy(dL, g->L, nv * 2, hipMemcpyHostToDevice));
        TN_HIP(hipMemcpy(dCnt, g->bl_cnt, nb * 4, hipMemcpyHostToDevice));
        TN_HIP(hipMemcpy(dBl, g->bl_lab, nb * TN_BL * 2, hipMemcpyHostToDevice));
        TN_HIP(hipMemcpy(dSlot, g->bl_slot, nb * 4, hipMemcpyHostToDevice));
        if (g->nslot > 0) {
            TN_HIP(hipMemcpy(dPhi, g->phi, (size_t)g->nslot * TN_SLOT * 4, hipMemcpyHostToDevice));
        }
        TN_HIP(hipMemcpy(dH, g->h, nv * 4, hipMemcpyHostToDevice));
        tnd::Field& f = p->f;
        f.nx = g->nx;
        f.ny = g->ny;
        f.nz = g->nz;
        f.nbx = g->nbx;
        f.nby = g->nby;
        f.nbz = g->nbz;
        f.vx = g->vs[0];
        f.vy = g->vs[1];
        f.vz = g->vs[2];
        f.L = dL;
        f.bl_cnt = dCnt;
        f.bl_lab = dBl;
        f.bl_slot = dSlot;
        f.phi = dPhi;
        f.h = dH;
        // node state
        p->dX0 = p->alloc<float4>(n);
        p->dX = p->alloc<float4>(n);
        p->dV = p->alloc<float4>(n);
        p->dF = p->alloc<float4>(n);
        p->dP0 = p->alloc<float4>(n);
        p->dPs = p->alloc<float4>(n);
        p->dTyp0 = p->alloc<uint8_t>(n);
        p->dTyp = p->alloc<uint8_t>(n);
        p->dPart0 = p->alloc<uint32_t>(n);
        p->dPart = p->alloc<uint32_t>(n);
        p->dLab = p->alloc<uint16_t>(n);
        p->dMv = p->alloc<float>(n);
        p->dPw = p->alloc<float>(n);
        p->dKey = p->alloc<int32_t>(n);
        p->dSorted = p->alloc<int32_t>(n);
        p->dNbrT = p->alloc<int32_t>((size_t)n * TN_K);
        p->dNnb = p->alloc<int32_t>(n);
        p->dOrder = p->alloc<int32_t>(n);
        p->dQueue = p->alloc<int32_t>(n);
        p->dCstart = p->alloc<int32_t>((size_t)p->nkeys + 1);
        for (int k = 0; k < 2; ++k) {
            p->dRk[k] = p->alloc<int32_t>(n);
            p->dRv[k] = p->alloc<int32_t>(n);
        }
        // radix sort geometry: 8-bit digits over the key range [0, nkeys)
        const int bits = p->nkeys <= 1 ? 0 : 32 - __builtin_clz((unsigned)(p->nkeys - 1));
        p->radix_passes = (bits + 7) / 8;
        p->radix_nblk = (int)std::max<unsigned>(1, blocks_for(n, tnk::kRadixTile));
        p->dBhist = p->alloc<int32_t>((size_t)256 * p->radix_nblk);
        const int64_t scan_max = std::max<int64_t>(p->nkeys + 1, (int64_t)256 * p->radix_nblk);
        p->dBsum = p->alloc<int32_t>(blocks_for(scan_max, tnk::kScanTile));
        p->nparts = (int)blocks_for(n, tnk::kBlock);
        p->dPartials = p->alloc<double2>(std::max(p->nparts, 1));
        p->dThr = p->alloc<float>(tnk::kHistThr);
        p->dCtl = p->alloc<tnk::Ctl>(1);
        TN_HIP(hipHostMalloc((void**)&p->hCtl, sizeof(tnk::Ctl), hipHostMallocDefault));
        std::memset(p->hCtl, 0, sizeof(tnk::Ctl));
        float thr[tnk::kHistThr];
        hist_thresholds(thr);
        TN_HIP(hipMemcpy(p->dThr, thr, sizeof thr, hipMemcpyHostToDevice));
        // the seeds
        if (n > 0) {
            float* dP3 = p->alloc<float>((size_t)3 * n);
            TN_HIP(hipMemcpy(dP3, seeds->P, sizeof(float) * 3 * n, hipMemcpyHostToDevice));
            k_init_nodes<<<blocks_for(n, 256), 256, 0, p->stream>>>(f, dP3, n, p->dX0);
            TN_HIP(hipGetLastError());
            std::vector<uint32_t> part(n);
            for (int i = 0; i < n; ++i) {
                part[i] = (uint32_t)seeds->part[2 * i] | ((uint32_t)seeds->part[2 * i + 1] << 16);
            }
            TN_HIP(hipMemcpy(p->dPart0, part.data(), sizeof(uint32_t) * n, hipMemcpyHostToDevice));
            TN_HIP(hipMemcpy(p->dTyp0, seeds->typ, n, hipMemcpyHostToDevice));
            TN_HIP(hipMemcpy(p->dLab, seeds->lab, sizeof(uint16_t) * n, hipMemcpyHostToDevice));
        }
        p->reset();
        *out = p.release();
    });
}

int tn_plan_reset(tn_plan* p) {
    return guarded([&] {
        arg_check(p != nullptr, "null plan");
        p->reset();
    });
}

int tn_plan_run(tn_plan* p, tn_relax_stats* st) {
    return guarded([&] {
        arg_check(p != nullptr, "null plan");
        p->reset();
        if (p->n > 0) {
            p->run_iters(p->prm.max_iters);
        }
        p->fill_stats(st);
    });
}

int tn_plan_step(tn_plan* p, int32_t iters, tn_relax_stats* st) {
    return guarded([&] {
        arg_check(p != nullptr && iters >= 0, "bad plan step");
        TN_HIP(hipSetDevice(p->device));
        if (p->n > 0) {
            p->run_iters(iters);
        }
        p->fill_stats(st);
    });
}

int tn_plan_download(tn_plan* p, tn_nodes_view* nd) {
    return guarded([&] {
        arg_check(p != nullptr, "null plan");
        check_nodes(nd);
        arg_check(nd->n == p->n, "nodes: count differs from the plan's");
        TN_HIP(hipSetDevice(p->device));
        TN_HIP(hipStreamSynchronize(p->stream));
        const int n = p->n;
        if (n == 0) {
            return;
        }
        std::vector<float4> X(n);
        std::vector<uint32_t> part(n);
        TN_HIP(hipMemcpy(X.data(), p->dX, sizeof(float4) * n, hipMemcpyDeviceToHost));
        TN_HIP(hipMemcpy(part.data(), p->dPart, sizeof(uint32_t) * n, hipMemcpyDeviceToHost));
        TN_HIP(hipMemcpy(nd->typ, p->dTyp, n, hipMemcpyDeviceToHost));
        for (int i = 0; i < n; ++i) {
            nd->P[3 * i] = X[i].x;
            nd->P[3 * i + 1] = X[i].y;
            nd->P[3 * i + 2] = X[i].z;
            nd->part[2 * i] = (uint16_t)(part[i] & 0xFFFF);
            nd->part[2 * i + 1] = (uint16_t)(part[i] >> 16);
        }
    });
}

int tn_plan_destroy(tn_plan* p) {
    return guarded([&] { delete p; });
}

int tn_plan_hash_info(const tn_plan* p, tn_hash_info* hi) {
    return guarded([&] {
        arg_check(p != nullptr && hi != nullptr, "null argument");
        std::memset(hi, 0, sizeof *hi);
        hi->b0 = p->H.b0;
        hi->ox = p->H.ox;
        hi->oy = p->H.oy;
        hi->oz = p->H.oz;
        hi->nlev = p->H.nlev;
        std::memcpy(hi->dim, p->H.dim, sizeof hi->dim);
        std::memcpy(hi->off, p->H.off, sizeof hi->off);
        hi->nkeys = (int32_t)p->nkeys;
    });
}

int tn_plan_rebuild(tn_plan* p, int32_t* key, int32_t* sorted, int32_t* cstart, int32_t* nbr, int32_t* nnb) {
    return guarded([&] {
        arg_check(p != nullptr, "null plan");
        TN_HIP(hipSetDevice(p->device));
        const int n = p->n;
        if (n > 0) {
            p->rebuild();
        }
        TN_HIP(hipStreamSynchronize(p->stream));
        if (n == 0) {
            if (cstart) {
                std::memset(cstart, 0, sizeof(int32_t) * (p->nkeys + 1));
            }
            return;
        }
        if (key) {
            TN_HIP(hipMemcpy(key, p->dKey, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
        }
        if (sorted) {
            TN_HIP(hipMemcpy(sorted, p->dSorted, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
        }
        if (cstart) {
            TN_HIP(hipMemcpy(cstart, p->dCstart, sizeof(int32_t) * (p->nkeys + 1), hipMemcpyDeviceToHost));
        }
        std::vector<int32_t> cnt(n);
        TN_HIP(hipMemcpy(cnt.data(), p->dNnb, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
        if (nnb) {
            std::memcpy(nnb, cnt.data(), sizeof(int32_t) * n);
        }
        if (nbr) {   // slot-major on the device, node-major (src/v2m_particles.cpp:764) here
            std::vector<int32_t> t((size_t)n * TN_K);
            TN_HIP(hipMemcpy(t.data(), p->dNbrT, sizeof(int32_t) * n * TN_K, hipMemcpyDeviceToHost));
            for (int i = 0; i < n; ++i) {
                for (int k = 0; k < TN_K; ++k) {
                    nbr[(size_t)i * TN_K + k] = k < cnt[i] ? t[(size_t)k * n + i] : -1;
                }
            }
        }
    });
}

int tn_plan_force(tn_plan* p, float* F) {
    return guarded([&] {
        arg_check(p != nullptr && F != nullptr, "null argument");
        TN_HIP(hipSetDevice(p->device));
        const int n = p->n;
        if (n == 0) {
            return;
        }
        tnk::k_force<<<blocks_for(n, tnk::kBlock), tnk::kBlock, 0, p->stream>>>(p->dX, p->dTyp, p->dNbrT, p->dNnb,
                                                                               p->prm.fscale, p->prm.fsurf, n, p->dF);
        TN_HIP(hipGetLastError());
        TN_HIP(hipStreamSynchronize(p->stream));
        TN_HIP(hipMemcpy(F, p->dF, sizeof(float4) * n, hipMemcpyDeviceToHost));
    });
}

int tn_relax(const tn_grid_view* g, const tn_relax_params* prm, tn_nodes_view* nd, tn_relax_stats* st,
             int32_t device) {
    tn_plan* p = nullptr;
    int rc = tn_plan_create(g, prm, nd, device, &p);
    if (rc != TN_OK) {
        return rc;
    }
    rc = tn_plan_run(p, st);
    if (rc == TN_OK) {
        rc = tn_plan_download(p, nd);
    }
    const std::string err = g_err;
    tn_plan_destroy(p);
    g_err = err;
    return rc;
}

}  // extern "C"
