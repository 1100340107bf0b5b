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
#include "tn_seed.hpp"

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
    // the K-nearest search may skip cells beyond a node's reach only if every node's
    // reach (t + skin) h fits the cell of its level, i.e. the largest h fits the top level
    const size_t nv = (size_t)g->nx * g->ny * g->nz;
    float hbig = 0.0f;
    for (size_t v = 0; v < nv; ++v) {
        hbig = std::max(hbig, g->h[v]);
    }
    H.prune = (prm->t + prm->skin) * hbig <= H.b0 * (float)(1 << (H.nlev - 1)) ? 1 : 0;
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

// Device buffers freed together (a plan's, or one seeding call's).
struct DevMem {
    std::vector<void*> owned;
    int device = -1;
    DevMem() = default;
    DevMem(const DevMem&) = delete;
    DevMem& operator=(const DevMem&) = delete;
    ~DevMem() {
        if (device >= 0) {
            (void)hipSetDevice(device);
        }
        for (void* p : owned) {
            (void)hipFree(p);
        }
    }
    template <class T>
    T* alloc(size_t count) {
        void* p = nullptr;
        TN_HIP(hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T)));
        owned.push_back(p);
        return static_cast<T*>(p);
    }
    template <class T>
    T* upload(const T* host, size_t count) {
        T* d = alloc<T>(count);
        if (count) {
            TN_HIP(hipMemcpy(d, host, count * sizeof(T), hipMemcpyHostToDevice));
        }
        return d;
    }
};

// the grid's arrays on the device (Grid, src/v2m_grid.h:45-81)
tnd::Field upload_grid(const tn_grid_view* g, DevMem& mem) {
    const size_t nv = (size_t)g->nx * g->ny * g->nz;
    const size_t nb = (size_t)g->nbx * g->nby * g->nbz;
    tnd::Field f;
    f.nx = g->nx;
    f.ny = g->ny;
    f.nz = g->nz;
    f.nbx = g->nbx;
    f.nby = g->nby;
    f.nbz = g->nbz;
    f.vx = g->vs[0];
    f.vy = g->vs[1];
    f.vz = g->vs[2];
    // a power of two whose reciprocal is a normal float: x / v == x * (1 / v) exactly
    auto pow2 = [](float v) {
        int e = 0;
        return std::isfinite(v) && v > 0.0f && std::frexp(v, &e) == 0.5f && e > -120 && e < 120;
    };
    f.pow2 = pow2(f.vx) && pow2(f.vy) && pow2(f.vz);
    f.ivx = 1.0f / f.vx;
    f.ivy = 1.0f / f.vy;
    f.ivz = 1.0f / f.vz;
    f.L = mem.upload(g->L, nv);
    f.bl_cnt = mem.upload(g->bl_cnt, nb);
    f.bl_lab = mem.upload(g->bl_lab, nb * TN_BL);
    f.bl_slot = mem.upload(g->bl_slot, nb);
    f.phi = mem.upload(g->phi, (size_t)g->nslot * TN_SLOT);
    f.h = mem.upload(g->h, nv);
    tnd::Field* self = mem.alloc<tnd::Field>(1);
    f.self = self;
    TN_HIP(hipMemcpy(self, &f, sizeof f, hipMemcpyHostToDevice));
    return f;
}

// Kernel classes of the per-kernel device timings.
enum KClass { KC_KEYS, KC_SCAN, KC_RADIX, KC_PERMUTE, KC_NEIGHBORS, KC_FORCE, KC_MOVE_EASY, KC_MOVE, KC_MOVE_REST,
              KC_STATS, KC_REFRESH, KC_CONTROL, KC_COUNT };
const char* const kKernelNames[KC_COUNT] = { "keys", "scan", "radix_sort", "permute", "neighbors",
                                             "force", "move_easy", "move", "move_rest", "stats", "refresh", "control" };

// Device time per kernel class from event pairs on one stream. Elapsed times are
// read only after the stream has been synchronised (collect), so timing adds
// event records to the stream but no host waits.
struct KernelTimer {
    bool on = false;
    hipStream_t stream = nullptr;
    std::vector<hipEvent_t> pool;
    struct Rec {
        int kc;
        hipEvent_t a, b;
    };
    std::vector<Rec> pending;
    double ms[KC_COUNT] = {};
    int64_t launches[KC_COUNT] = {};
    int64_t items[KC_COUNT] = {};

    ~KernelTimer() {
        for (const Rec& r : pending) {
            pool.push_back(r.a);
            pool.push_back(r.b);
        }
        for (hipEvent_t e : pool) {
            (void)hipEventDestroy(e);
        }
    }
    hipEvent_t take() {
        if (pool.empty()) {
            hipEvent_t e;
            TN_HIP(hipEventCreate(&e));
            return e;
        }
        hipEvent_t e = pool.back();
        pool.pop_back();
        return e;
    }
    hipEvent_t begin() {
        if (!on) {
            return nullptr;
        }
        hipEvent_t e = take();
        TN_HIP(hipEventRecord(e, stream));
        return e;
    }
    void end(int kc, hipEvent_t a) {
        if (!on || !a) {
            return;
        }
        hipEvent_t b = take();
        TN_HIP(hipEventRecord(b, stream));
        pending.push_back({ kc, a, b });
    }
    void collect() {   // the stream has been synchronised
        for (const Rec& r : pending) {
            float t = 0;
            TN_HIP(hipEventElapsedTime(&t, r.a, r.b));
            ms[r.kc] += t;
            launches[r.kc] += 1;
            pool.push_back(r.a);
            pool.push_back(r.b);
        }
        pending.clear();
    }
    void count(int kc, int64_t m) {   // nodes (or cells) a timed launch processed
        if (on) {
            items[kc] += m;
        }
    }
    void clear() {
        std::fill(ms, ms + KC_COUNT, 0.0);
        std::fill(launches, launches + KC_COUNT, 0);
        std::fill(items, items + KC_COUNT, 0);
    }
};

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
    KernelTimer timer;
    // device buffers
    DevMem mem;
    // node state by slot (renumbered in cell order at every rebuild; *b: the other
    // buffer of each pair), and the seeds by id
    float4 *dX0 = nullptr, *dX = nullptr, *dV = nullptr, *dF = nullptr, *dP0 = nullptr, *dPs = nullptr;
    float4 *dXb = nullptr, *dVb = nullptr;
    int2* dMs = nullptr;
    uint8_t *dTyp0 = nullptr, *dTyp = nullptr, *dTypb = nullptr;
    uint32_t *dPart0 = nullptr, *dPart = nullptr, *dPartb = nullptr;
    uint16_t *dLab0 = nullptr, *dLab = nullptr, *dLabb = nullptr;
    int32_t *dId = nullptr, *dIdb = nullptr, *dSlotOf = nullptr;
    const int32_t* dKeySorted = nullptr;   // the keys in slot order after the last rebuild
    float *dMv = nullptr, *dPw = nullptr, *dThr = nullptr;
    int32_t *dKey = nullptr, *dCstart = nullptr, *dSorted = nullptr, *dNbrT = nullptr, *dNnb = nullptr;
    int32_t *dQueue = nullptr, *dMoveQ = nullptr, *dMoveQ2 = nullptr, *dBhist = nullptr, *dBsum = nullptr;
    uint8_t* dFlag = nullptr;   // nodes past a full skin this iteration
    int32_t* dRk[2] = { nullptr, nullptr };
    int32_t* dRv[2] = { nullptr, nullptr };
    double2* dPartials = nullptr;
    tnk::Ctl* dCtl = nullptr;
    tnk::Ctl* hCtl = nullptr;   // pinned
    int nparts = 0, radix_passes = 0, radix_nblk = 0, refresh_blocks = 1;
    // host-side loop state
    int rebuilds = 0;
    bool finished = false;

    template <class T>
    T* alloc(size_t count) {
        return mem.alloc<T>(count);
    }

    ~tn_plan() {
        if (device >= 0) {
            (void)hipSetDevice(device);
        }
        if (stream) {
            (void)hipStreamSynchronize(stream);
        }
        if (hCtl) {
            (void)hipHostFree(hCtl);
        }
        if (stream) {
            (void)hipStreamDestroy(stream);
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
        in.Ps = dPs;
        in.Ms = dMs;
        in.slot_of = dSlotOf;
        in.t = prm.t;
        in.skin = prm.skin;
        return in;
    }

    // the rebuild lambda (src/v2m_particles.cpp:771-820) and the type order
    // (src/v2m_gpu.cpp:288-302)
    void rebuild() {
        const unsigned nb = blocks_for(n, tnk::kBlock);
        hipEvent_t t = timer.begin();
        tnk::k_fill_i32<<<blocks_for(nkeys + 1, tnk::kBlock), tnk::kBlock, 0, stream>>>(dCstart, nkeys + 1, 0);
        tnk::k_keys<<<nb, tnk::kBlock, 0, stream>>>(H, dX, n, tskin, dKey, dCstart);
        timer.end(KC_KEYS, t);
        t = timer.begin();
        scan(dCstart, dCstart, nkeys + 1);
        timer.end(KC_SCAN, t);
        t = timer.begin();
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
        timer.end(KC_RADIX, t);
        t = timer.begin();
        dKeySorted = radix_passes == 0 ? dKey : dRk[(radix_passes - 1) & 1];
        const tnk::NodeArrays cur{ dX, dV, dTyp, dPart, dLab, dId }, nxt{ dXb, dVb, dTypb, dPartb, dLabb, dIdb };
        tnk::k_permute<<<nb, tnk::kBlock, 0, stream>>>(dSorted, n, cur, nxt, dPs, dP0, dMs, dSlotOf);
        std::swap(dX, dXb);
        std::swap(dV, dVb);
        std::swap(dTyp, dTypb);
        std::swap(dPart, dPartb);
        std::swap(dLab, dLabb);
        std::swap(dId, dIdb);
        timer.end(KC_PERMUTE, t);
        t = timer.begin();
        tnk::k_truss<<<blocks_for(n, tnk::kTrussWave), tnk::kTrussWave, 0, stream>>>(H, truss_in(), n, dNbrT, dNnb);
        timer.end(KC_NEIGHBORS, t);
        TN_HIP(hipGetLastError());
        for (int kc : { KC_KEYS, KC_RADIX, KC_PERMUTE, KC_NEIGHBORS }) {
            timer.count(kc, n);
        }
        timer.count(KC_SCAN, nkeys + 1);
        ++rebuilds;
    }

    // device state back to the seeds, then the first rebuild
    void reset() {
        TN_HIP(hipSetDevice(device));
        TN_HIP(hipMemcpyAsync(dX, dX0, sizeof(float4) * n, hipMemcpyDeviceToDevice, stream));
        TN_HIP(hipMemcpyAsync(dTyp, dTyp0, n, hipMemcpyDeviceToDevice, stream));
        TN_HIP(hipMemcpyAsync(dPart, dPart0, sizeof(uint32_t) * n, hipMemcpyDeviceToDevice, stream));
        TN_HIP(hipMemcpyAsync(dLab, dLab0, sizeof(uint16_t) * n, hipMemcpyDeviceToDevice, stream));
        if (n > 0) {   // slot = id
            k_iota<<<blocks_for(n, tnk::kBlock), tnk::kBlock, 0, stream>>>(dId, n);
        }
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
        if (n > 0) {
            rebuild();
        }
        TN_HIP(hipStreamSynchronize(stream));
        timer.collect();
    }

    // one iteration (src/v2m_particles.cpp:828-946); returns false once converged
    bool iterate() {
        const unsigned nb = blocks_for(n, tnk::kBlock);
        tnk::k_iter_begin<<<1, 64, 0, stream>>>(dCtl);
        hipEvent_t t = timer.begin();
        tnk::k_force<<<nb, tnk::kBlock, 0, stream>>>(dX, dTyp, dNbrT, dNnb, prm.fscale, prm.fsurf, n, dF);
        timer.end(KC_FORCE, t);
        t = timer.begin();
        tnk::MoveArgs ma{ prm.dt, prm.maxstep, prm.snap, fire };
        tnk::k_move_easy<<<nb, tnk::kBlock, 0, stream>>>(f, ma, dCtl, n, dF, dX, dV, dTyp, dPart, dLab, dMv,
                                                         dPw, dMoveQ, dMoveQ2);
        timer.end(KC_MOVE_EASY, t);
        t = timer.begin();
        const unsigned g1 = std::min<unsigned>(nb, tnk::kMoveGrid), g2 = std::min<unsigned>(nb, tnk::kMoveRestGrid);
        tnk::k_move<true><<<g1, tnk::kBlock, 0, stream>>>(f, ma, dCtl, dMoveQ, dF, dX, dV, dTyp, dPart, dLab, dMv, dPw,
                                                          dMoveQ2);
        timer.end(KC_MOVE, t);
        t = timer.begin();
        tnk::k_move<false><<<g2, tnk::kBlock, 0, stream>>>(f, ma, dCtl, dMoveQ2, dF, dX, dV, dTyp, dPart, dLab, dMv,
                                                           dPw, nullptr);
        timer.end(KC_MOVE_REST, t);
        t = timer.begin();
        tnk::k_stats<<<nparts, tnk::kBlock, 0, stream>>>(dX, dP0, dMv, dPw, n, prm.skin, fire, dThr, dCtl, dFlag,
                                                         dPartials);
        timer.end(KC_STATS, t);
        t = timer.begin();
        tnk::k_refresh_list<<<nb, tnk::kBlock, 0, stream>>>(dFlag, n, dCtl, dQueue);
        tnk::k_truss_node<<<refresh_blocks, tnk::kWave, 0, stream>>>(H, truss_in(), n, dQueue, dCtl, dNbrT, dNnb, dP0);
        timer.end(KC_REFRESH, t);
        t = timer.begin();
        tnk::CtlArgs ca{ n, prm.max_iters, fire, rebuild_div, prm.dptol, nparts };
        tnk::k_control<<<1, tnk::kBlock, 0, stream>>>(ca, dCtl, dPartials, dPw, dSlotOf);
        timer.end(KC_CONTROL, t);
        TN_HIP(hipGetLastError());
        TN_HIP(hipMemcpyAsync(hCtl, dCtl, sizeof(tnk::Ctl), hipMemcpyDeviceToHost, stream));
        TN_HIP(hipStreamSynchronize(stream));
        timer.collect();
        for (int kc : { KC_FORCE, KC_MOVE_EASY, KC_STATS }) {
            timer.count(kc, n);
        }
        timer.count(KC_MOVE, hCtl->nmove);
        timer.count(KC_MOVE_REST, hCtl->nmove_rest);
        timer.count(KC_REFRESH, hCtl->nrefresh);
        timer.count(KC_CONTROL, 1);
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
        timer.collect();
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
        const double* ms = timer.ms;
        st->ms_hash = ms[KC_KEYS] + ms[KC_SCAN] + ms[KC_RADIX] + ms[KC_PERMUTE] + ms[KC_NEIGHBORS];
        st->ms_force = ms[KC_FORCE];
        st->ms_move = ms[KC_MOVE_EASY] + ms[KC_MOVE] + ms[KC_MOVE_REST] + ms[KC_STATS] + ms[KC_REFRESH] + ms[KC_CONTROL];
        if (n > 0) {   // type counts (src/v2m_particles.cpp:968-975), counted on the device
            tnk::k_order_begin<<<1, 64, 0, stream>>>(dCtl);
            tnk::k_order_count<<<blocks_for(n, tnk::kBlock), tnk::kBlock, 0, stream>>>(dTyp, n, dCtl);
            TN_HIP(hipGetLastError());
            tnk::Ctl c2;
            TN_HIP(hipMemcpyAsync(hCtl, dCtl, sizeof(tnk::Ctl), hipMemcpyDeviceToHost, stream));
            TN_HIP(hipStreamSynchronize(stream));
            c2 = *hCtl;
            st->n_interface = c2.order_cnt[0];
            st->n_junction = c2.order_cnt[1];
            st->n_interior = c2.order_cnt[2];
            st->n_corner = c2.order_cnt[3];
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

// The seeding of seed_cpu_body (src/v2m_particles.cpp:637-669): per-voxel counts,
// their scan, the fill in voxel order and the classification, on one device.
// The seeds stay in `mem`; returns their count.
struct Seeds {
    int64_t n = 0;
    float* P = nullptr;
    uint16_t* lab = nullptr;
    uint8_t* typ = nullptr;
    uint16_t* part = nullptr;
};

Seeds run_seeding(const tn_grid_view* g, const uint8_t* grade, const tn_relax_params* prm, int dev, DevMem& mem,
                  bool fill) {
    check_grid(g);
    arg_check(prm != nullptr && grade != nullptr, "seeding: null argument");
    if (prm->jseed > 0.0f || prm->corners) {
        throw TnError(TN_ERR_UNSUPPORTED, "seeding: junction-line seeds (jseed > 0) and corner nodes are not "
                                          "generated on the device; set jseed = 0 and corners = 0");
    }
    arg_check(prm->nseed >= 1 && prm->nseed <= 64, "seeding: nseed out of range");
    TN_HIP(hipSetDevice(dev));
    mem.device = dev;
    const tnd::Field f = upload_grid(g, mem);
    const int64_t nv = (int64_t)g->nx * g->ny * g->nz;
    // v2m_seed_spacing (src/opencl/v2m_seed_body.cl:445-447), with the host's powf
    std::vector<float> spacing(prm->nseed);
    for (int L = 0; L < prm->nseed; ++L) {
        spacing[L] = prm->nseed <= 1 ? g->hmin
                                     : g->hmin * std::pow(g->hmax / g->hmin, (float)L / (float)(prm->nseed - 1));
        arg_check(spacing[L] > 0.0f && std::isfinite(spacing[L]), "seeding: bad seed spacing");
    }
    tnk::SeedArgs sa;
    sa.grade = mem.upload(grade, (size_t)nv);
    sa.spacing = mem.upload(spacing.data(), spacing.size());
    sa.nseed = prm->nseed;
    int32_t* offs = mem.alloc<int32_t>((size_t)nv + 1);
    int32_t* bsum = mem.alloc<int32_t>(blocks_for(nv + 1, tnk::kScanTile));
    tnk::k_fill_i32<<<blocks_for(nv + 1, tnk::kBlock), tnk::kBlock>>>(offs, nv + 1, 0);
    tnk::k_seed_count<<<blocks_for(nv, tnk::kBlock), tnk::kBlock>>>(f, sa, offs);
    const unsigned nb = blocks_for(nv + 1, tnk::kScanTile);
    tnk::k_scan_reduce<<<nb, tnk::kBlock>>>(offs, nv + 1, bsum);
    tnk::k_scan_top<<<1, tnk::kBlock>>>(bsum, nb);
    tnk::k_scan_down<<<nb, tnk::kBlock>>>(offs, offs, nv + 1, bsum);
    TN_HIP(hipGetLastError());
    int32_t total = 0;
    TN_HIP(hipMemcpy(&total, offs + nv, sizeof total, hipMemcpyDeviceToHost));
    Seeds s;
    s.n = total;
    if (!fill || total == 0) {
        return s;
    }
    s.P = mem.alloc<float>((size_t)3 * total);
    s.lab = mem.alloc<uint16_t>(total);
    s.typ = mem.alloc<uint8_t>(total);
    s.part = mem.alloc<uint16_t>((size_t)2 * total);
    tnk::k_seed_fill<<<blocks_for(nv, tnk::kBlock), tnk::kBlock>>>(f, sa, offs, s.P, s.lab);
    tnk::k_seed_classify<<<blocks_for(total, tnk::kBlock), tnk::kBlock>>>(f, total, s.P, s.lab, s.typ, s.part);
    TN_HIP(hipGetLastError());
    TN_HIP(hipDeviceSynchronize());
    return s;
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
        p->H = make_hash(g, prm, &p->nkeys);
        TN_HIP(hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking));
        p->timer.stream = p->stream;
        p->timer.on = prm->verbose || std::getenv("TN_KERNEL_TIMING") != nullptr;
        const int n = p->n;
        p->mem.device = dev;
        p->f = upload_grid(g, p->mem);
        const tnd::Field& f = p->f;
        // node state
        p->dX0 = p->alloc<float4>(n);
        p->dX = p->alloc<float4>(n);
        p->dXb = p->alloc<float4>(n);
        p->dV = p->alloc<float4>(n);
        p->dVb = p->alloc<float4>(n);
        p->dF = p->alloc<float4>(n);
        p->dP0 = p->alloc<float4>(n);
        p->dPs = p->alloc<float4>(n);
        p->dMs = p->alloc<int2>(n);
        p->dTyp0 = p->alloc<uint8_t>(n);
        p->dTyp = p->alloc<uint8_t>(n);
        p->dTypb = p->alloc<uint8_t>(n);
        p->dPart0 = p->alloc<uint32_t>(n);
        p->dPart = p->alloc<uint32_t>(n);
        p->dPartb = p->alloc<uint32_t>(n);
        p->dLab0 = p->alloc<uint16_t>(n);
        p->dLab = p->alloc<uint16_t>(n);
        p->dLabb = p->alloc<uint16_t>(n);
        p->dId = p->alloc<int32_t>(n);
        p->dIdb = p->alloc<int32_t>(n);
        p->dSlotOf = p->alloc<int32_t>(n);
        p->dMv = p->alloc<float>(n);
        p->dPw = p->alloc<float>(n);
        p->dKey = p->alloc<int32_t>(n);
        p->dSorted = p->alloc<int32_t>(n);
        p->dNbrT = p->alloc<int32_t>((size_t)n * TN_K);
        p->dNnb = p->alloc<int32_t>(n);
        p->dQueue = p->alloc<int32_t>(n);
        p->dMoveQ = p->alloc<int32_t>(n);
        p->dMoveQ2 = p->alloc<int32_t>(n);
        p->dFlag = p->alloc<uint8_t>(n);
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
        p->nparts = (int)std::max<unsigned>(1, std::min<unsigned>(blocks_for(n, tnk::kBlock), tnk::kStatsBlocks));
        p->refresh_blocks = (int)std::max<int64_t>(1, std::min<int64_t>(n, 2048));   // one wave per listed node
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
            TN_HIP(hipMemcpy(p->dLab0, seeds->lab, sizeof(uint16_t) * n, hipMemcpyHostToDevice));
        }
        p->reset();
        *out = p.release();
    });
}

int tn_plan_set_timing(tn_plan* p, int32_t on) {
    return guarded([&] {
        arg_check(p != nullptr, "null plan");
        TN_HIP(hipSetDevice(p->device));
        TN_HIP(hipStreamSynchronize(p->stream));
        p->timer.collect();
        p->timer.clear();
        p->timer.on = on != 0;
    });
}

int tn_plan_timings(const tn_plan* p, tn_kernel_time* out, int32_t cap, int32_t* count) {
    return guarded([&] {
        arg_check(p != nullptr && count != nullptr && cap >= 0 && (cap == 0 || out != nullptr), "bad timings call");
        *count = KC_COUNT;
        for (int k = 0; k < KC_COUNT && k < cap; ++k) {
            std::memset(&out[k], 0, sizeof out[k]);
            std::strncpy(out[k].name, kKernelNames[k], sizeof out[k].name - 1);
            out[k].launches = p->timer.launches[k];
            out[k].ms = p->timer.ms[k];
            out[k].items = p->timer.items[k];
        }
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
        std::vector<uint8_t> typ(n);
        std::vector<int32_t> id(n);
        TN_HIP(hipMemcpy(X.data(), p->dX, sizeof(float4) * n, hipMemcpyDeviceToHost));
        TN_HIP(hipMemcpy(part.data(), p->dPart, sizeof(uint32_t) * n, hipMemcpyDeviceToHost));
        TN_HIP(hipMemcpy(typ.data(), p->dTyp, n, hipMemcpyDeviceToHost));
        TN_HIP(hipMemcpy(id.data(), p->dId, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
        for (int t = 0; t < n; ++t) {   // slot t holds node id[t]
            const int i = id[t];
            nd->P[3 * i] = X[t].x;
            nd->P[3 * i + 1] = X[t].y;
            nd->P[3 * i + 2] = X[t].z;
            nd->typ[i] = typ[t];
            nd->part[2 * i] = (uint16_t)(part[t] & 0xFFFF);
            nd->part[2 * i + 1] = (uint16_t)(part[t] >> 16);
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
        // by node id, as the reference's rebuild leaves them (src/v2m_particles.cpp:771-820)
        std::vector<int32_t> id(n), ks(n), cnt(n);
        TN_HIP(hipMemcpy(id.data(), p->dId, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
        TN_HIP(hipMemcpy(ks.data(), p->dKeySorted, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
        TN_HIP(hipMemcpy(cnt.data(), p->dNnb, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
        std::vector<int32_t> kid(n);
        for (int t = 0; t < n; ++t) {
            kid[id[t]] = ks[t];
        }
        if (key) {
            std::memcpy(key, kid.data(), sizeof(int32_t) * n);
        }
        if (sorted) {   // the counting sort's order: by key, ascending id within a cell
            std::vector<int32_t> o(n);
            for (int i = 0; i < n; ++i) {
                o[i] = i;
            }
            std::stable_sort(o.begin(), o.end(), [&](int32_t a, int32_t b) { return kid[a] < kid[b]; });
            std::memcpy(sorted, o.data(), sizeof(int32_t) * n);
        }
        if (cstart) {
            TN_HIP(hipMemcpy(cstart, p->dCstart, sizeof(int32_t) * (p->nkeys + 1), hipMemcpyDeviceToHost));
        }
        if (nnb) {
            for (int t = 0; t < n; ++t) {
                nnb[id[t]] = cnt[t];
            }
        }
        if (nbr) {   // slot-major by slot on the device, node-major by id (src/v2m_particles.cpp:764) here
            std::vector<int32_t> tr((size_t)n * TN_K);
            TN_HIP(hipMemcpy(tr.data(), p->dNbrT, sizeof(int32_t) * n * TN_K, hipMemcpyDeviceToHost));
            for (int t = 0; t < n; ++t) {
                const size_t i = (size_t)id[t];
                for (int k = 0; k < TN_K; ++k) {
                    nbr[i * TN_K + k] = k < cnt[t] ? id[tr[(size_t)k * n + t]] : -1;
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
        std::vector<float4> Fs(n);
        std::vector<int32_t> id(n);
        TN_HIP(hipMemcpy(Fs.data(), p->dF, sizeof(float4) * n, hipMemcpyDeviceToHost));
        TN_HIP(hipMemcpy(id.data(), p->dId, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
        for (int t = 0; t < n; ++t) {   // by node id
            std::memcpy(F + 4 * (size_t)id[t], &Fs[t], sizeof(float4));
        }
    });
}

int tn_seed_count(const tn_grid_view* g, const uint8_t* grade, const tn_relax_params* prm, int32_t device,
                  int64_t* n) {
    return guarded([&] {
        arg_check(n != nullptr, "null count");
        DevMem mem;
        *n = run_seeding(g, grade, prm, pick_device(device), mem, false).n;
    });
}

int tn_seed(const tn_grid_view* g, const uint8_t* grade, const tn_relax_params* prm, int32_t device, int64_t n,
            float* P, uint16_t* lab, uint8_t* typ, uint16_t* part) {
    return guarded([&] {
        arg_check(n >= 0 && (n == 0 || (P && lab && typ && part)), "seeding: null output array");
        DevMem mem;
        const Seeds s = run_seeding(g, grade, prm, pick_device(device), mem, true);
        arg_check(s.n == n, "seeding: the output holds " + std::to_string(n) + " nodes, the seeding made " +
                                std::to_string(s.n) + " (tn_seed_count)");
        if (s.n == 0) {
            return;
        }
        TN_HIP(hipMemcpy(P, s.P, sizeof(float) * 3 * s.n, hipMemcpyDeviceToHost));
        TN_HIP(hipMemcpy(lab, s.lab, sizeof(uint16_t) * s.n, hipMemcpyDeviceToHost));
        TN_HIP(hipMemcpy(typ, s.typ, s.n, hipMemcpyDeviceToHost));
        TN_HIP(hipMemcpy(part, s.part, sizeof(uint16_t) * 2 * s.n, hipMemcpyDeviceToHost));
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
