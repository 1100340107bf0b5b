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
#include <array>
#include <map>
#include <thread>
#include <type_traits>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "tn_kernels.hpp"
#include "tn_seed.hpp"
#include "tn_thin.hpp"

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
// (hbig: the largest h of the grid)
tnd::Hash make_hash_geom(int nx, int ny, int nz, const float* vs, float hmin, float hmax, float hbig,
                         const tn_relax_params* prm, int64_t* nkeys_out) {
    tnd::Hash H;
    std::memset(&H, 0, sizeof H);
    H.b0 = (prm->t + prm->skin) * hmin;
    H.ox = -0.5f * vs[0];
    H.oy = -0.5f * vs[1];
    H.oz = -0.5f * vs[2];
    const float ext[3] = { (float)nx * vs[0], (float)ny * vs[1], (float)nz * vs[2] };
    H.nlev = 1;
    while (H.nlev < TN_MAXLEV && H.b0 * (float)(1 << (H.nlev - 1)) < (prm->t + prm->skin) * hmax) {
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
    H.prune = (prm->t + prm->skin) * hbig <= H.b0 * (float)(1 << (H.nlev - 1)) ? 1 : 0;
    return H;
}

tnd::Hash make_hash(const tn_grid_view* g, const tn_relax_params* prm, int64_t* nkeys_out) {
    const size_t nv = (size_t)g->nx * g->ny * g->nz;
    float hbig = 0.0f;
    for (size_t v = 0; v < nv; ++v) {
        hbig = std::max(hbig, g->h[v]);
    }
    return make_hash_geom(g->nx, g->ny, g->nz, g->vs, g->hmin, g->hmax, hbig, prm, nkeys_out);
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

// Freed device buffers kept for reuse, by device and size, so that creating and
// destroying plans (a batch of 64 volumes, a loop of single calls) rarely calls
// hipMalloc / hipFree, each of which costs tens of microseconds to milliseconds
// and a device-wide sync. At most TN_MEM_CACHE_MB (default 8192; 0: off) are kept;
// the buffers are never returned to the driver before the process ends.
struct MemCache {
    std::mutex mu;
    std::multimap<std::pair<int, size_t>, void*> free;
    size_t bytes = 0, cap = 0;
    MemCache() {
        const char* e = std::getenv("TN_MEM_CACHE_MB");
        cap = (e && e[0] ? (size_t)std::strtoull(e, nullptr, 10) : (size_t)8192) << 20;
    }
    void* take(int dev, size_t b) {
        std::lock_guard<std::mutex> g(mu);
        const auto it = free.find({ dev, b });
        if (it == free.end()) {
            return nullptr;
        }
        void* p = it->second;
        free.erase(it);
        bytes -= b;
        return p;
    }
    bool give(int dev, size_t b, void* p) {
        std::lock_guard<std::mutex> g(mu);
        if (bytes + b > cap) {
            return false;
        }
        free.emplace(std::make_pair(dev, b), p);
        bytes += b;
        return true;
    }
};

MemCache& mem_cache() {
    static MemCache* c = new MemCache;   // (never destroyed: the runtime may be gone at exit)
    return *c;
}

// Device memory owned by one object and freed with it (a plan's, or one seeding
// call's). Small buffers are carved from shared 16 MB chunks (256-byte aligned): a
// plan holds a few dozen buffers. Buffers come from and go back to the MemCache; the
// device is synchronised before they go back, as hipFree would. With an arena
// (set_arena), every buffer is carved from that caller-given device memory instead
// and nothing is allocated or freed; a request beyond its end throws.
struct DevMem {
    static constexpr size_t kChunk = (size_t)16 << 20, kAlign = 256;
    struct Owned {
        void* p;
        size_t bytes;
        int dev;
    };
    std::vector<Owned> owned;
    char* chunk = nullptr;
    size_t chunk_left = 0;
    int device = -1;
    char* arena = nullptr;
    size_t arena_cap = 0, arena_used = 0;
    DevMem() = default;
    DevMem(const DevMem&) = delete;
    DevMem& operator=(const DevMem&) = delete;
    ~DevMem() {
        if (owned.empty()) {
            return;
        }
        if (device >= 0) {
            (void)hipSetDevice(device);
        }
        (void)hipDeviceSynchronize();   // no work in flight may still use them
        for (const Owned& o : owned) {
            if (!mem_cache().give(o.dev, o.bytes, o.p)) {
                (void)hipFree(o.p);
            }
        }
    }
    // a buffer of exactly `bytes` on the current device: a cached one, else a new one
    void* fresh(size_t bytes) {
        int dev = 0;
        TN_HIP(hipGetDevice(&dev));
        void* p = mem_cache().take(dev, bytes);
        if (!p) {
            TN_HIP(hipMalloc(&p, bytes));
        }
        owned.push_back({ p, bytes, dev });
        return p;
    }
    void set_arena(void* base, size_t cap) {
        arena = static_cast<char*>(base);
        arena_cap = cap;
        arena_used = 0;
    }
    void* raw(size_t bytes) {
        bytes = (std::max<size_t>(bytes, 1) + kAlign - 1) / kAlign * kAlign;
        if (arena) {
            if (bytes > arena_cap - arena_used) {
                throw TnError(TN_ERR_ARG, "scratch memory: this relaxation needs more than the " +
                                              std::to_string(arena_cap >> 20) + " MiB the no-allocation entry "
                                              "point holds (" + std::to_string(arena_used + bytes) +
                                              " bytes requested so far)");
            }
            void* p = arena + arena_used;
            arena_used += bytes;
            return p;
        }
        if (bytes > kChunk / 4) {   // a buffer of its own
            return fresh(bytes);
        }
        if (bytes > chunk_left) {
            chunk = static_cast<char*>(fresh(kChunk));
            chunk_left = kChunk;
        }
        void* p = chunk;
        chunk += bytes;
        chunk_left -= bytes;
        return p;
    }
    template <class T>
    T* alloc(size_t count) {
        return static_cast<T*>(raw(count * sizeof(T)));
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

// the exact voxel-size reciprocals of a Field: a power of two whose reciprocal is a
// normal float has x / v == x * (1 / v) exactly
void set_reciprocals(tnd::Field& f) {
    auto pow2 = [](float v) {
        int e = 0;
        return std::isfinite(v) && v > 0.0f && std::frexp(v, &e) == 0.5f && e > -120 && e < 120;
    };
    f.pow2 = pow2(f.vx) && pow2(f.vy) && pow2(f.vz);
    f.ivx = 1.0f / f.vx;
    f.ivy = 1.0f / f.vy;
    f.ivz = 1.0f / f.vz;
}

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
    set_reciprocals(f);
    f.L = mem.upload(g->L, nv);
    f.bl_cnt = mem.upload(g->bl_cnt, nb);
    f.bl_lab = mem.upload(g->bl_lab, nb * TN_BL);
    f.bl_slot = mem.upload(g->bl_slot, nb);
    f.phi = mem.upload(g->phi, (size_t)g->nslot * TN_SLOT);
    f.h = mem.upload(g->h, nv);
    f.gm = g->gm;
    f.gI = f.gm > 0 ? mem.upload(g->gI, nv) : nullptr;
    f.gTW = f.gm > 0 ? mem.upload(g->gTW, 2 * (size_t)f.gm) : (f.gm < 0 ? mem.upload(g->gTW, (size_t)g->ngTW) : nullptr);
    tnd::Field* self = mem.alloc<tnd::Field>(1);
    f.self = self;
    TN_HIP(hipMemcpy(self, &f, sizeof f, hipMemcpyHostToDevice));
    return f;
}

// Kernel classes of the per-kernel device timings.
enum KClass { KC_KEYS, KC_SCAN, KC_RADIX, KC_PERMUTE, KC_NEIGHBORS, KC_FORCE, KC_MOVE_EASY, KC_MOVE, KC_MOVE_INT,
              KC_MOVE_REST, KC_MOVE_DEF,
              KC_STATS, KC_REFRESH, KC_CONTROL, KC_COUNT };
constexpr uint32_t kAllClasses = (1u << KC_COUNT) - 1u;

const char* const kKernelNames[KC_COUNT] = { "keys", "scan", "radix_sort", "permute", "neighbors",
                                             "force", "move_easy", "move", "move_int", "move_rest", "move_def",
                                             "stats", "refresh", "control" };

// Device time per kernel class from event pairs on the stream of each launch.
// Elapsed times are read only after the streams have been synchronised (collect),
// so timing adds event records to the streams but no host waits; `mask` selects the
// classes timed (bit k: class k), so that a few can be timed at a small cost.
struct KernelTimer {
    uint32_t mask = 0;
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
    hipEvent_t begin(int kc, hipStream_t s = nullptr) {
        if (!((mask >> kc) & 1u)) {
            return nullptr;
        }
        hipEvent_t e = take();
        TN_HIP(hipEventRecord(e, s ? s : stream));
        return e;
    }
    void end(int kc, hipEvent_t a, hipStream_t s = nullptr) {
        if (!a) {
            return;
        }
        hipEvent_t b = take();
        TN_HIP(hipEventRecord(b, s ? s : stream));
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
        if ((mask >> kc) & 1u) {
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
    int force_serial = 0;   // TN_SERIAL_POWER: the serial power sum every FIRE iteration
    int rest_lanes = tnk::kRestLanes;   // the general move pass's nodes per wave (a launch shape)
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
    // nodes a lean move pass has handed back to the general pass in an earlier iteration
    // of this run (by slot): k_force queues them for the general pass directly
    uint8_t *dSticky = nullptr, *dStickyb = nullptr;
    const int32_t* dKeySorted = nullptr;   // the keys in slot order after the last rebuild
    tnk::BrickMap bricks{};                // the K-nearest search's query order (brick positions)
    int32_t* dBrick = nullptr;             // [bricks.size + 1]: node counts, then starts, by brick position
    float *dMv = nullptr, *dPw = nullptr, *dThr = nullptr;
    int32_t *dKey = nullptr, *dCstart = nullptr, *dSorted = nullptr, *dNbrT = nullptr, *dNnb = nullptr;
    int32_t *dQueue = nullptr, *dBhist = nullptr, *dBsum = nullptr;
    // the move queues: interface (lean), interior (lean), general, handed back
    int32_t *dMoveQ = nullptr, *dMoveQi = nullptr, *dMoveQ2 = nullptr, *dMoveQd = nullptr;
    uint8_t* dFlag = nullptr;   // nodes past a full skin this iteration
    int32_t* dRk[2] = { nullptr, nullptr };
    int32_t* dRv[2] = { nullptr, nullptr };
    double2* dPartials = nullptr;
    tnk::Ctl* dCtl = nullptr;
    tnk::Ctl* hCtl = nullptr;   // pinned (pageable: hctl_local, in a plan that allocates nothing)
    tnk::Ctl hctl_local{};
    // a plan that allocates nothing: its buffers are carved from a caller-given arena,
    // it runs on the null stream with no side streams, events or pinned memory
    bool noalloc = false;
    hipEvent_t ev_iter = nullptr;   // behind each iteration's control-block copy
    // the general pass over the junction nodes runs on side stream 0 from the forces on
    // (ev_force), beside k_move_easy and the lean interface pass; the lean interior pass
    // runs on side stream 1 (ev_easy); ev_rest and ev_int join them back
    hipStream_t side[2] = { nullptr, nullptr };
    hipEvent_t ev_force = nullptr, ev_easy = nullptr, ev_rest = nullptr, ev_int = nullptr;
    // false: the three passes run one after another on the plan's stream (many plans
    // at once keep the device busy by themselves; cross-stream waits of many plans
    // would queue behind each other on the few hardware queues)
    bool overlap_moves = true;
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
        if (stream || noalloc) {
            (void)hipStreamSynchronize(stream);
        }
        if (hCtl && !noalloc) {
            (void)hipHostFree(hCtl);
        }
        for (hipStream_t s : side) {
            if (s) {
                (void)hipStreamSynchronize(s);
            }
        }
        for (hipEvent_t e : { ev_iter, ev_force, ev_easy, ev_rest, ev_int }) {
            if (e) {
                (void)hipEventDestroy(e);
            }
        }
        for (hipStream_t s : side) {
            if (s) {
                (void)hipStreamDestroy(s);
            }
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
        in.order = dSorted;   // (after a rebuild's permutation: the query order)
        in.t = prm.t;
        in.skin = prm.skin;
        return in;
    }

    // the rebuild lambda (src/v2m_particles.cpp:771-820) and the type order
    // (src/v2m_gpu.cpp:288-302)
    void rebuild() {
        const unsigned nb = blocks_for(n, tnk::kBlock);
        hipEvent_t t = timer.begin(KC_KEYS);
        tnk::k_fill_i32<<<blocks_for(nkeys + 1, tnk::kBlock), tnk::kBlock, 0, stream>>>(dCstart, nkeys + 1, 0);
        tnk::k_keys<<<nb, tnk::kBlock, 0, stream>>>(H, dX, n, tskin, dKey, dCstart);
        timer.end(KC_KEYS, t);
        t = timer.begin(KC_SCAN);
        scan(dCstart, dCstart, nkeys + 1);
        timer.end(KC_SCAN, t);
        t = timer.begin(KC_RADIX);
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
        t = timer.begin(KC_PERMUTE);
        dKeySorted = radix_passes == 0 ? dKey : dRk[(radix_passes - 1) & 1];
        const tnk::NodeArrays cur{ dX, dV, dTyp, dPart, dLab, dId, dSticky },
            nxt{ dXb, dVb, dTypb, dPartb, dLabb, dIdb, dStickyb };
        tnk::k_permute<<<nb, tnk::kBlock, 0, stream>>>(dSorted, n, cur, nxt, dPs, dP0, dMs, dSlotOf);
        std::swap(dX, dXb);
        std::swap(dV, dVb);
        std::swap(dTyp, dTypb);
        std::swap(dPart, dPartb);
        std::swap(dLab, dLabb);
        std::swap(dId, dIdb);
        std::swap(dSticky, dStickyb);
        timer.end(KC_PERMUTE, t);
        t = timer.begin(KC_NEIGHBORS);
        // the query order, into dSorted (free once the nodes are permuted)
        tnk::k_fill_i32<<<blocks_for(bricks.size + 1, tnk::kBlock), tnk::kBlock, 0, stream>>>(dBrick, bricks.size + 1,
                                                                                             0);
        tnk::k_brick_count<<<blocks_for(nkeys, tnk::kBlock), tnk::kBlock, 0, stream>>>(H, bricks, dCstart, dBrick);
        scan(dBrick, dBrick, (int64_t)bricks.size + 1);
        tnk::k_brick_order<<<nb, tnk::kBlock, 0, stream>>>(H, bricks, dKeySorted, n, dCstart, dBrick, dSorted);
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
        if (n > 0) {   // slot = id; no node handed back yet (the buffers are padded to 4 bytes)
            k_iota<<<blocks_for(n, tnk::kBlock), tnk::kBlock, 0, stream>>>(dId, n);
            const int words = (n + 3) / 4;
            tnk::k_fill_i32<<<blocks_for(words, tnk::kBlock), tnk::kBlock, 0, stream>>>(
                reinterpret_cast<int32_t*>(dSticky), words, 0);
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

    // one iteration (src/v2m_particles.cpp:828-946), enqueued: the kernels, then the
    // control block's copy to the host and ev_iter behind them
    void enqueue_iteration() {
        const unsigned nb = blocks_for(n, tnk::kBlock);
        tnk::k_iter_begin<<<1, 64, 0, stream>>>(dCtl);
        hipEvent_t t = timer.begin(KC_FORCE);
        // (the junction nodes queued for the general pass, unless the move is the voxel trap's)
        tnk::k_force<<<nb, tnk::kBlock, 0, stream>>>(dX, dTyp, dNbrT, dNnb, prm.fscale, prm.fsurf, n, dF, dCtl,
                                                     prm.voxel_trap ? nullptr : dMoveQ2, dFlag, dSticky);
        timer.end(KC_FORCE, t);
        t = timer.begin(KC_MOVE_EASY);
        tnk::MoveArgs ma{ prm.dt, prm.maxstep, prm.snap, fire, rest_lanes };
        if (prm.voxel_trap) {   // the staircase move: a lane per node, then a wave per node that scans
            tnk::k_move_voxel<<<nb, tnk::kBlock, 0, stream>>>(f, ma, dCtl, n, dF, dX, dV, dTyp, dPart, dLab, dMv,
                                                              dPw, dMoveQ, &dCtl->nmove);
            timer.end(KC_MOVE_EASY, t);
            t = timer.begin(KC_MOVE);
            tnk::k_move_voxel_wave<<<std::min<unsigned>(nb, tnk::kMoveGrid), tnk::kBlock, 0, stream>>>(
                f, ma, dCtl, dMoveQ, &dCtl->nmove, dF, dX, dV, dTyp, dPart, dLab, dMv, dPw);
            timer.end(KC_MOVE, t);
            enqueue_iteration_tail(nb);
            return;
        }
        // (label volumes: the move kernels compiled without the other field modes)
        auto moves = [&](auto plain) {
            constexpr bool kP = decltype(plain)::value;
            // The queued nodes: the junction nodes' general pass starts on a side stream as
            // soon as the forces are done, beside k_move_easy; then the two lean passes at
            // once over the queues k_move_easy fills, each bounded by its slowest node, then
            // the general pass over the nodes they hand back.
            const unsigned g1 = std::min<unsigned>(nb, tnk::kMoveGrid), g2 = std::min<unsigned>(nb, tnk::kMoveRestGrid);
            const hipStream_t s0 = overlap_moves ? side[0] : stream, s1 = overlap_moves ? side[1] : stream;
            if (overlap_moves) {
                TN_HIP(hipEventRecord(ev_force, stream));
                TN_HIP(hipStreamWaitEvent(s0, ev_force, 0));
            }
            hipEvent_t tr = timer.begin(KC_MOVE_REST, s0);
            tnk::k_move<tnd::kGeneral, kP><<<g2, tnk::kBlock, 0, s0>>>(f, ma, dCtl, dMoveQ2, dF, dX, dV, dTyp, dPart, dLab, dMv,
                                                                   dPw, &dCtl->nmove_rest, nullptr, nullptr, nullptr);
            timer.end(KC_MOVE_REST, tr, s0);
            if (overlap_moves) {
                TN_HIP(hipEventRecord(ev_rest, s0));
            }
            tnk::k_move_easy<kP><<<nb, tnk::kBlock, 0, stream>>>(f, ma, dCtl, n, dF, dX, dV, dTyp, dPart, dLab, dMv,
                                                             dPw, dMoveQ, dMoveQi, dFlag);
            timer.end(KC_MOVE_EASY, t);
            if (overlap_moves) {
                TN_HIP(hipEventRecord(ev_easy, stream));
                TN_HIP(hipStreamWaitEvent(s1, ev_easy, 0));
            }
            t = timer.begin(KC_MOVE_INT, s1);
            tnk::k_move<tnd::kLeanInterior, kP><<<g1, tnk::kBlock, 0, s1>>>(f, ma, dCtl, dMoveQi, dF, dX, dV, dTyp, dPart, dLab,
                                                                        dMv, dPw, &dCtl->nmove_int, &dCtl->nmove_def,
                                                                        dMoveQd, dSticky);
            timer.end(KC_MOVE_INT, t, s1);
            if (overlap_moves) {
                TN_HIP(hipEventRecord(ev_int, s1));
            }
            t = timer.begin(KC_MOVE);
            tnk::k_move<tnd::kLeanInterface, kP><<<g1, tnk::kBlock, 0, stream>>>(f, ma, dCtl, dMoveQ, dF, dX, dV, dTyp, dPart,
                                                                             dLab, dMv, dPw, &dCtl->nmove, &dCtl->nmove_def,
                                                                             dMoveQd, dSticky);
            timer.end(KC_MOVE, t);
            if (overlap_moves) {
                TN_HIP(hipStreamWaitEvent(stream, ev_int, 0));
            }
            t = timer.begin(KC_MOVE_DEF);
            tnk::k_move<tnd::kGeneral, kP><<<g2, tnk::kBlock, 0, stream>>>(f, ma, dCtl, dMoveQd, dF, dX, dV, dTyp, dPart, dLab,
                                                                       dMv, dPw, &dCtl->nmove_def, nullptr, nullptr, nullptr);
            timer.end(KC_MOVE_DEF, t);
            if (overlap_moves) {
                TN_HIP(hipStreamWaitEvent(stream, ev_rest, 0));
            }
        };
        if (f.gm == 0) {
            moves(std::true_type{});
        } else {
            moves(std::false_type{});
        }
        enqueue_iteration_tail(nb);
    }

    // stats, own-list refreshes, control, and the control block's copy to the host
    void enqueue_iteration_tail(unsigned nb) {
        hipEvent_t t = timer.begin(KC_STATS);
        tnk::k_stats<<<nparts, tnk::kBlock, 0, stream>>>(dX, dP0, dMv, dPw, n, prm.skin, fire, dThr, dCtl, dFlag,
                                                         dPartials);
        timer.end(KC_STATS, t);
        // the control before the own-list refreshes, which it does not read: they are
        // skipped when a rebuild or the loop's end follows
        t = timer.begin(KC_CONTROL);
        tnk::CtlArgs ca{ n, prm.max_iters, fire, rebuild_div, prm.dptol, nparts, force_serial };
        tnk::k_control<<<1, tnk::kBlock, 0, stream>>>(ca, dCtl, dPartials, dPw, dSlotOf);
        timer.end(KC_CONTROL, t);
        t = timer.begin(KC_REFRESH);
        tnk::k_refresh_list<<<nb, tnk::kBlock, 0, stream>>>(dFlag, n, dCtl, dQueue);
        tnk::k_truss_node<<<refresh_blocks, tnk::kWave, 0, stream>>>(H, truss_in(), n, dQueue, dCtl, dNbrT, dNnb, dP0);
        timer.end(KC_REFRESH, t);
        TN_HIP(hipGetLastError());
        TN_HIP(hipMemcpyAsync(hCtl, dCtl, sizeof(tnk::Ctl), hipMemcpyDeviceToHost, stream));
        if (ev_iter) {
            TN_HIP(hipEventRecord(ev_iter, stream));
        }
    }

    // after ev_iter: the host's part of the iteration (the decisions are the device's,
    // in the control block). Enqueues the rebuild when one is due; false once the
    // relaxation has converged.
    bool finish_iteration() {
        timer.collect();
        for (int kc : { KC_FORCE, KC_MOVE_EASY, KC_STATS }) {
            timer.count(kc, n);
        }
        timer.count(KC_MOVE, hCtl->nmove);
        timer.count(KC_MOVE_INT, hCtl->nmove_int);
        timer.count(KC_MOVE_REST, hCtl->nmove_rest);
        timer.count(KC_MOVE_DEF, hCtl->nmove_def);
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

    // the arguments of enqueue_iteration's launches, for the grouped launches of many
    // plans (they change at every rebuild, which swaps the node buffers)
    tnk::PlanIter iter_args() const {
        tnk::PlanIter a{};
        a.f = f;
        a.H = H;
        a.tin = truss_in();
        a.ma = tnk::MoveArgs{ prm.dt, prm.maxstep, prm.snap, fire, rest_lanes };
        a.ca = tnk::CtlArgs{ n, prm.max_iters, fire, rebuild_div, prm.dptol, nparts, force_serial };
        a.n = n;
        a.fire = fire;
        a.nparts = nparts;
        a.voxel = prm.voxel_trap ? 1 : 0;
        a.fscale = prm.fscale;
        a.fsurf = prm.fsurf;
        a.skin = prm.skin;
        a.ctl = dCtl;
        a.nbrT = dNbrT;
        a.nnb = dNnb;
        a.X = dX;
        a.V = dV;
        a.F = dF;
        a.P0 = dP0;
        a.typ = dTyp;
        a.part = dPart;
        a.lab = dLab;
        a.mv = dMv;
        a.pw = dPw;
        a.thr = dThr;
        a.flag = dFlag;
        a.partials = dPartials;
        a.rqueue = dQueue;
        a.mq = dMoveQ;
        a.mqi = dMoveQi;
        a.mq2 = dMoveQ2;
        a.mqd = dMoveQd;
        a.slot_of = dSlotOf;
        a.sticky = dSticky;
        const int nb = (int)blocks_for(n, tnk::kBlock);
        a.nblk[tnk::kShNodes] = nb;
        a.nblk[tnk::kShLean] = std::min(nb, tnk::kMoveGrid);
        a.nblk[tnk::kShGen] = std::min(nb, tnk::kMoveRestGrid);
        a.nblk[tnk::kShParts] = nparts;
        a.nblk[tnk::kShRefresh] = refresh_blocks;
        a.nblk[tnk::kShOne] = 1;
        return a;
    }

    bool iterate() {
        enqueue_iteration();
        if (ev_iter) {
            TN_HIP(hipEventSynchronize(ev_iter));
        } else {
            TN_HIP(hipStreamSynchronize(stream));
        }
        return finish_iteration();
    }

    bool more() const {   // iterations remain
        return n > 0 && !finished && hCtl->it < prm.max_iters;
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
        st->power_fallbacks = c.power_fallbacks;
        const double* ms = timer.ms;
        st->ms_hash = ms[KC_KEYS] + ms[KC_SCAN] + ms[KC_RADIX] + ms[KC_PERMUTE] + ms[KC_NEIGHBORS];
        st->ms_force = ms[KC_FORCE];
        st->ms_move = ms[KC_MOVE_EASY] + ms[KC_MOVE] + ms[KC_MOVE_INT] + ms[KC_MOVE_REST] + ms[KC_MOVE_DEF] + ms[KC_STATS] + ms[KC_REFRESH] + ms[KC_CONTROL];
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

// A plan on device dev: make_grid(mem) puts the grid into the plan's memory and
// returns its Field; H and nkeys are the hash geometry. With an arena, the plan
// allocates nothing (DevMem::set_arena; tn_plan::noalloc).
template <class MakeGrid>
tn_plan* plan_build(const tn_relax_params* prm, const tn_nodes_view* seeds, int dev, const tnd::Hash& H, int64_t nkeys,
                    MakeGrid make_grid, void* arena = nullptr, size_t arena_cap = 0) {
        TN_HIP(hipSetDevice(dev));
        std::unique_ptr<tn_plan> p(new tn_plan);
        p->device = dev;
        p->prm = *prm;
        p->n = (int)seeds->n;
        p->fire = prm->fire && !prm->voxel_trap ? 1 : 0;
        p->rebuild_div = prm->rebuild_div > 0 ? prm->rebuild_div : 300;
        p->tskin = prm->t + prm->skin;
        p->H = H;
        p->nkeys = nkeys;
        p->noalloc = arena != nullptr;
        if (p->noalloc) {   // the null stream, the passes one after another, no timing
            p->overlap_moves = false;
            p->mem.set_arena(arena, arena_cap);
        } else {
            TN_HIP(hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking));
            TN_HIP(hipEventCreateWithFlags(&p->ev_iter, hipEventDisableTiming));
            for (hipStream_t& s : p->side) {
                TN_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
            }
            for (hipEvent_t* e : { &p->ev_force, &p->ev_easy, &p->ev_rest, &p->ev_int }) {
                TN_HIP(hipEventCreateWithFlags(e, hipEventDisableTiming));
            }
        }
        p->timer.stream = p->stream;
        p->timer.mask = !p->noalloc && (prm->verbose || std::getenv("TN_KERNEL_TIMING") != nullptr) ? kAllClasses : 0u;
        {
            const char* e = std::getenv("TN_SERIAL_POWER");
            p->force_serial = e && e[0] && e[0] != '0' ? 1 : 0;
        }
        const int n = p->n;
        p->mem.device = dev;
        p->f = make_grid(p->mem);
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
        p->dSticky = p->alloc<uint8_t>(n);
        p->dStickyb = p->alloc<uint8_t>(n);
        p->dSlotOf = p->alloc<int32_t>(n);
        p->dMv = p->alloc<float>(n);
        p->dPw = p->alloc<float>(n);
        p->dKey = p->alloc<int32_t>(n);
        p->dSorted = p->alloc<int32_t>(n);
        p->dNbrT = p->alloc<int32_t>((size_t)n * TN_K);
        p->dNnb = p->alloc<int32_t>(n);
        p->dQueue = p->alloc<int32_t>(n);
        p->dMoveQ = p->alloc<int32_t>(n);
        p->dMoveQi = p->alloc<int32_t>(n);
        p->dMoveQ2 = p->alloc<int32_t>(n);
        p->dMoveQd = p->alloc<int32_t>(n);
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
        {   // the query order's brick positions
            int64_t pos = 0;
            for (int L = 0; L < p->H.nlev; ++L) {
                int64_t nbl = 1;
                for (int a = 0; a < 3; ++a) {
                    p->bricks.nb[L][a] = (p->H.dim[L][a] + tnk::kBrickEdge - 1) / tnk::kBrickEdge;
                    nbl *= p->bricks.nb[L][a];
                }
                p->bricks.off[L] = (int)pos;
                pos += nbl * tnk::kBrickEdge * tnk::kBrickEdge * tnk::kBrickEdge;
            }
            arg_check(pos < ((int64_t)1 << 31) - 1, "hash: too many cells for the query order");
            p->bricks.size = (int)pos;
            p->dBrick = p->alloc<int32_t>((size_t)pos + 1);
        }
        const int64_t scan_max = std::max<int64_t>(std::max<int64_t>(p->nkeys, p->bricks.size) + 1,
                                                   (int64_t)256 * p->radix_nblk);
        p->dBsum = p->alloc<int32_t>(blocks_for(scan_max, tnk::kScanTile));
        p->nparts = (int)std::max<unsigned>(1, std::min<unsigned>(blocks_for(n, tnk::kBlock), tnk::kStatsBlocks));
        p->refresh_blocks = (int)std::max<int64_t>(1, std::min<int64_t>(n, 2048));   // one wave per listed node
        p->dPartials = p->alloc<double2>(std::max(p->nparts, 1));
        p->dThr = p->alloc<float>(tnk::kHistThr);
        p->dCtl = p->alloc<tnk::Ctl>(1);
        if (p->noalloc) {
            p->hCtl = &p->hctl_local;
        } else {
            TN_HIP(hipHostMalloc((void**)&p->hCtl, sizeof(tnk::Ctl), hipHostMallocDefault));
        }
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
        return p.release();
}

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

// A shape program (the layout of src/opencl/v2m_sdf_body.cl:1-45) whose every read
// the interpreter can make lies inside its n values: the header, the brick table,
// and each label's code walked op by op to its end, with every bounding-box skip
// landing on an op of the same code.
void check_program(const float* prog, int64_t n) {
    auto at = [&](int64_t i) -> double {
        arg_check(i >= 0 && i < n, "shape program: a read beyond its " + std::to_string(n) + " values");
        return (double)prog[i];
    };
    auto idx = [&](int64_t i) -> int64_t {
        const double v = at(i);
        arg_check(v >= 0.0 && v < (double)n && v == std::floor(v), "shape program: a bad offset");
        return (int64_t)v;
    };
    arg_check(n >= 9, "shape program: too short");
    const double dN = at(0);
    arg_check(dN >= 1.0 && dN <= 65535.0 && dN == std::floor(dN), "shape program: bad label count");
    const int N = (int)dN;
    at(8 + N);
    auto code = [&](int64_t start) {
        std::vector<int64_t> ops, skips;
        int64_t pc = start;
        for (int guard = 0;; ++guard) {
            arg_check(guard < 100000, "shape program: a code without its end");
            const int op = (int)at(pc);
            ops.push_back(pc);
            if (op == tnd::sdf::kEnd) {
                break;
            }
            if (op == tnd::sdf::kPrim) {
                const int type = (int)at(pc + 1);
                const int np = tnd::sdf::nparam(type);
                at(pc + 1 + std::max(np, 1));
                pc += 2 + np;
            } else if (op == tnd::sdf::kBBox) {
                const double len = at(pc + 7);
                at(pc + 8);
                arg_check(len >= 0.0 && len == std::floor(len), "shape program: a bad skip");
                skips.push_back(pc + 9 + (int64_t)len);
                pc += 9;
            } else if (op == tnd::sdf::kMConst || op == tnd::sdf::kConst || op == tnd::sdf::kScale) {
                at(pc + 1);
                pc += 2;
            } else {
                arg_check(tnd::sdf::is_minmax(op) || op == tnd::sdf::kNeg || op == tnd::sdf::kAdd,
                          "shape program: an unknown op");
                pc += 1;
            }
        }
        for (int64_t t : skips) {
            arg_check(std::binary_search(ops.begin(), ops.end(), t), "shape program: a skip off the code");
        }
    };
    for (int l = 0; l < N; ++l) {
        code(idx(5 + l));
    }
    const int64_t T = (int64_t)at(8 + N);
    if (T > 0) {
        const int64_t nbx = (int64_t)at(T), nby = (int64_t)at(T + 1), nbz = (int64_t)at(T + 2);
        arg_check(nbx >= 1 && nby >= 1 && nbz >= 1 && at(T + 3) > 0.0, "shape program: a bad brick table");
        at(T + 3 + nbx * nby * nbz);
        for (int64_t b = 0; b < nbx * nby * nbz; ++b) {
            const int64_t e = idx(T + 4 + b);
            for (int l = 0; l < N; ++l) {
                code(idx(e + l));
            }
        }
    }
}

void check_grid(const tn_grid_view* g) {
    arg_check(g != nullptr, "null grid");
    arg_check(g->gm <= 0 || (g->gI && g->gTW && g->ngTW >= 2 * (int64_t)g->gm),
              "grid: a gray-scale grid (gm > 0) needs gI and 2 gm values in gTW");
    if (g->gm < 0) {
        arg_check(g->gTW && g->ngTW >= 9, "grid: a shape grid (gm < 0) needs its program in gTW");
        check_program(g->gTW, g->ngTW);
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

// The seeding of seed_cpu_body (src/v2m_particles.cpp:637-726) on one device: the
// lattice (per-voxel counts, their scan, the fill in voxel order, the
// classification), then the junction-line seeds (seed_junctions,
// src/v2m_particles.cpp:80-244) and the corner nodes. The seeds stay in `mem`.
struct Seeds {
    int64_t n = 0;
    float* P = nullptr;
    uint16_t* lab = nullptr;
    uint8_t* typ = nullptr;
    uint16_t* part = nullptr;
    uint16_t* part3 = nullptr;
};

// Device scans and sorts of the seeding, on the null stream.
struct SeedOps {
    DevMem& mem;
    // offs[0..m) holds counts; afterwards offs[0..m] their exclusive scan. Returns the total.
    int32_t scan(int32_t* offs, int64_t m) {
        const unsigned nb = blocks_for(m + 1, tnk::kScanTile);
        int32_t* bsum = mem.alloc<int32_t>(nb);
        tnk::k_scan_reduce<<<nb, tnk::kBlock>>>(offs, m + 1, bsum);
        tnk::k_scan_top<<<1, tnk::kBlock>>>(bsum, nb);
        tnk::k_scan_down<<<nb, tnk::kBlock>>>(offs, offs, m + 1, bsum);
        TN_HIP(hipGetLastError());
        int32_t total = 0;
        TN_HIP(hipMemcpy(&total, offs + m, sizeof total, hipMemcpyDeviceToHost));
        return total;
    }
    // a zeroed count array of m + 1 entries
    int32_t* counts(int64_t m) {
        int32_t* c = mem.alloc<int32_t>((size_t)m + 1);
        tnk::k_fill_i32<<<blocks_for(m + 1, tnk::kBlock), tnk::kBlock>>>(c, m + 1, 0);
        return c;
    }
    // Stable sort of (keys, vals) by keys in [0, 2^bits): LSD passes of 8 bits.
    // vals == nullptr: the values are the element indices. The results are left in
    // the returned pair (keys, vals); the inputs are overwritten.
    std::pair<int32_t*, int32_t*> sort(int32_t* keys, int32_t* vals, int n, int bits) {
        const int passes = (bits + 7) / 8;
        const int nblk = (int)std::max<unsigned>(1, blocks_for(n, tnk::kRadixTile));
        int32_t* bhist = mem.alloc<int32_t>((size_t)256 * nblk);
        int32_t* k2 = mem.alloc<int32_t>(n);
        int32_t* v2 = mem.alloc<int32_t>(n);
        if (vals == nullptr) {
            vals = mem.alloc<int32_t>(n);
            k_iota<<<blocks_for(n, tnk::kBlock), tnk::kBlock>>>(vals, n);
        }
        for (int p = 0; p < passes; ++p) {
            tnk::k_radix_hist<<<nblk, tnk::kBlock>>>(keys, n, 8 * p, nblk, bhist);
            scan_exclusive(bhist, (int64_t)256 * nblk);
            tnk::k_radix_scatter<<<nblk, tnk::kBlock>>>(keys, vals, n, 8 * p, nblk, bhist, k2, v2);
            std::swap(keys, k2);
            std::swap(vals, v2);
        }
        TN_HIP(hipGetLastError());
        return { keys, vals };
    }
    void scan_exclusive(int32_t* a, int64_t m) {   // in place, no total
        const unsigned nb = blocks_for(m, tnk::kScanTile);
        int32_t* bsum = mem.alloc<int32_t>(nb);
        tnk::k_scan_reduce<<<nb, tnk::kBlock>>>(a, m, bsum);
        tnk::k_scan_top<<<1, tnk::kBlock>>>(bsum, nb);
        tnk::k_scan_down<<<nb, tnk::kBlock>>>(a, a, m, bsum);
    }
};

int bit_length(uint32_t x) {
    int b = 0;
    while (x) {
        ++b;
        x >>= 1;
    }
    return b;
}

// seed_junctions (src/v2m_particles.cpp:80-244): the candidates at the 3-label
// vertices, one kept per (level, cell, label triple) -- the smallest vertex -- in
// key order, and the lattice nodes within 0.4 h of a kept seed dropped. Writes the
// result into `out` (capacity: the lattice count + the candidate count + extra)
// and returns its node count.
int64_t junction_pass(float ghmax, const tnd::Field& f, const tnk::SeedArgs& sa, float cellfac,
                      const tnk::VertexGrid& vg, const Seeds& lat, int64_t extra, SeedOps& ops, Seeds& out) {
    DevMem& mem = ops.mem;
    const int n = (int)lat.n;
    int32_t* flag = ops.counts(vg.nvert);
    tnk::k_vertex_mark<0><<<blocks_for(vg.nvert, tnk::kBlock), tnk::kBlock>>>(f, vg, flag);
    const int m3 = ops.scan(flag, vg.nvert);
    tnk::JunctionCands jc{};
    jc.m = m3;
    int nk = 0;
    tnk::KeptSeeds ks{};
    int32_t *sorted_bkey = nullptr, *sidx = nullptr;
    if (m3 > 0) {
        int32_t* list = mem.alloc<int32_t>(m3);
        tnk::k_compact_index<<<blocks_for(vg.nvert, tnk::kBlock), tnk::kBlock>>>(flag, vg.nvert, list);
        jc.x = mem.alloc<float>((size_t)3 * m3);
        jc.lab3 = mem.alloc<int32_t>((size_t)3 * m3);
        jc.key = mem.alloc<int32_t>((size_t)7 * m3);
        jc.ok = mem.alloc<int32_t>((size_t)m3 + 1);
        tnk::k_fill_i32<<<blocks_for(m3 + 1, tnk::kBlock), tnk::kBlock>>>(jc.ok, m3 + 1, 0);
        tnk::k_junction_cand<<<blocks_for(m3, tnk::kBlock), tnk::kBlock>>>(f, sa, cellfac, vg, list, jc);
        const int nc = ops.scan(jc.ok, m3);
        if (nc > 0) {
            // the candidates (in vertex order), compacted
            int32_t* cl = mem.alloc<int32_t>(nc);
            tnk::k_compact_index<<<blocks_for(m3, tnk::kBlock), tnk::kBlock>>>(jc.ok, m3, cl);
            tnk::JunctionCands cc{};
            cc.m = nc;
            cc.x = mem.alloc<float>((size_t)3 * nc);
            cc.lab3 = mem.alloc<int32_t>((size_t)3 * nc);
            cc.key = mem.alloc<int32_t>((size_t)7 * nc);
            tnk::k_gather_cands<<<blocks_for(nc, tnk::kBlock), tnk::kBlock>>>(jc, cl, cc);
            // each key field's range, for the LSD passes
            int32_t* range = mem.alloc<int32_t>(14);
            tnk::k_key_range_init<<<1, 64>>>(range);
            tnk::k_key_range<<<blocks_for(nc, tnk::kBlock), tnk::kBlock>>>(cc.key, nc, range);
            int32_t hr[14];
            TN_HIP(hipMemcpy(hr, range, sizeof hr, hipMemcpyDeviceToHost));
            // stable passes from the least significant field up: ties stay in vertex order
            int32_t* perm = mem.alloc<int32_t>(nc);
            k_iota<<<blocks_for(nc, tnk::kBlock), tnk::kBlock>>>(perm, nc);
            int32_t* fk = mem.alloc<int32_t>(nc);
            for (int e = 6; e >= 0; --e) {
                const uint32_t span = (uint32_t)((int64_t)hr[2 * e + 1] - hr[2 * e]);
                if (span == 0) {
                    continue;
                }
                tnk::k_gather_field<<<blocks_for(nc, tnk::kBlock), tnk::kBlock>>>(cc.key + (size_t)e * nc, perm, nc,
                                                                                 -hr[2 * e], fk);
                const auto r = ops.sort(fk, perm, nc, bit_length(span));
                fk = r.first;
                perm = r.second;
            }
            int32_t* head = ops.counts(nc);
            tnk::k_key_heads<<<blocks_for(nc, tnk::kBlock), tnk::kBlock>>>(cc.key, perm, nc, head);
            nk = ops.scan(head, nc);
            // the drop hash of the kept seeds (src/v2m_particles.cpp:160-181)
            ks.nk = nk;
            ks.cell = 0.4f * ghmax;
            ks.ox = f.vx;
            ks.oy = f.vy;
            ks.oz = f.vz;
            const float ext[3] = { (f.nx + 1) * f.vx, (f.ny + 1) * f.vy, (f.nz + 1) * f.vz };
            int64_t cells = 1;
            for (int e = 0; e < 3; ++e) {
                ks.hd[e] = (int)(ext[e] / ks.cell) + 2;
                cells *= ks.hd[e];
            }
            if (cells >= ((int64_t)1 << 31)) {
                throw TnError(TN_ERR_UNSUPPORTED, "seeding: the junction-seed hash needs more than 2^31 cells "
                                                  "(hmax too small for the grid)");
            }
            ks.x = mem.alloc<float>((size_t)3 * nk);
            ks.lab3 = mem.alloc<int32_t>((size_t)3 * nk);
            ks.bkey = mem.alloc<int32_t>(nk);
            tnk::k_keep_seeds<<<blocks_for(nc, tnk::kBlock), tnk::kBlock>>>(cc, perm, head, ks);
            int32_t* bk = mem.alloc<int32_t>(nk);
            TN_HIP(hipMemcpy(bk, ks.bkey, sizeof(int32_t) * nk, hipMemcpyDeviceToDevice));
            const auto r = ops.sort(bk, nullptr, nk, bit_length((uint32_t)(cells - 1)));
            sorted_bkey = r.first;
            sidx = r.second;
        }
    }
    int32_t* keep = ops.counts(n);
    int n1 = n;
    if (n > 0) {
        if (nk > 0) {
            tnk::k_drop_near_seeds<<<blocks_for(n, tnk::kBlock), tnk::kBlock>>>(f, n, lat.P, lat.typ, ks, sorted_bkey,
                                                                              sidx, keep);
        } else {
            tnk::k_fill_i32<<<blocks_for(n, tnk::kBlock), tnk::kBlock>>>(keep, n, 1);
        }
        n1 = ops.scan(keep, n);
    }
    const int64_t total = (int64_t)n1 + nk;
    out.n = total;
    const size_t cap = (size_t)std::max<int64_t>(total + extra, 1);
    out.P = mem.alloc<float>(3 * cap);
    out.lab = mem.alloc<uint16_t>(cap);
    out.typ = mem.alloc<uint8_t>(cap);
    out.part = mem.alloc<uint16_t>(2 * cap);
    out.part3 = mem.alloc<uint16_t>(cap);
    if (n + nk > 0) {
        tnk::SeedNodes in{ lat.P, lat.lab, lat.typ, lat.part, lat.part3 };
        tnk::SeedNodes o{ out.P, out.lab, out.typ, out.part, out.part3 };
        tnk::k_merge_junctions<<<blocks_for((int64_t)n + nk, tnk::kBlock), tnk::kBlock>>>(in, n, keep, ks, o);
        TN_HIP(hipGetLastError());
    }
    return total;
}

Seeds seeding_core(const tnd::Field& f, const uint8_t* d_grade, float hmin, float hmax, const tn_relax_params* prm,
                   DevMem& mem);

// v2m_h_at (src/opencl/v2m_particle_body.cl:30-36) on the host's copy of the sizing
float host_h_at(const tn_grid_view* g, const float* p) {
    int i = (int)std::floor(p[0] / g->vs[0] + 0.5f), j = (int)std::floor(p[1] / g->vs[1] + 0.5f),
        k = (int)std::floor(p[2] / g->vs[2] + 0.5f);
    i = i < 0 ? 0 : (i >= g->nx ? g->nx - 1 : i);
    j = j < 0 ? 0 : (j >= g->ny ? g->ny - 1 : j);
    k = k < 0 ? 0 : (k >= g->nz ? g->nz - 1 : k);
    return g->h[(size_t)i + (size_t)g->nx * ((size_t)j + (size_t)g->ny * k)];
}

// add_feature_nodes (src/v2m_particles.cpp:304-620): pinned CORNER nodes on the shape
// input's sharp features (g->feat: corners, segments, circles, polylines), spaced by
// the sizing along each curve and kept where the labels round them differ; the seeds
// within 0.4 h of one are dropped and the kept ones appended. The curve geometry, the
// bisections' bookkeeping, the label counts and the greedy spacing run on the host
// (a few thousand points, the reference's float and double arithmetic); the label
// queries and the seeds' drop test run on the device.
Seeds add_feature_nodes(const tn_grid_view* g, const tnd::Field& f, const Seeds& s, DevMem& mem) {
    SeedOps ops{ mem };
    auto hat = [&](const float* p) { return host_h_at(g, p); };
    // the masks of a batch of points (k_feat_mask), one device round
    auto masks = [&](const std::vector<float>& X) {
        const int m = (int)(X.size() / 3);
        std::vector<unsigned long long> out(m);
        if (m == 0) {
            return out;
        }
        DevMem tmp;
        tmp.device = mem.device;
        const float* dX = tmp.upload(X.data(), X.size());
        unsigned long long* dm = tmp.alloc<unsigned long long>(m);
        tnk::k_feat_mask<<<blocks_for(m, tnk::kBlock), tnk::kBlock>>>(f, dX, m, dm);
        TN_HIP(hipGetLastError());
        TN_HIP(hipMemcpy(out.data(), dm, sizeof(unsigned long long) * m, hipMemcpyDeviceToHost));
        return out;
    };
    constexpr int nsamp = 256;
    struct Curve {
        int type;
        const float* q;
        float u[3], w[3];
        std::vector<float> pts;
        void at(float t, float* o) const {   // the curve at parameter t in [0, 1]
            if (type == 4) {   // a polyline: [n, points], by its points' index
                const int n = (int)q[0];
                const float uu = t * (float)(n - 1);
                const int i = std::min(n - 2, std::max(0, (int)uu));
                const float fr = uu - (float)i;
                for (int a = 0; a < 3; ++a) {
                    o[a] = q[1 + 3 * i + a] + fr * (q[1 + 3 * (i + 1) + a] - q[1 + 3 * i + a]);
                }
            } else if (type == 2) {
                for (int a = 0; a < 3; ++a) {
                    o[a] = q[a] + t * (q[3 + a] - q[a]);
                }
            } else {
                const float th = 6.28318531f * t, c = std::cos(th), sn = std::sin(th);
                for (int a = 0; a < 3; ++a) {
                    o[a] = q[a] + q[6] * (c * u[a] + sn * w[a]);
                }
            }
        }
    };
    std::vector<float> fp, tp, rp;   // candidates: corners, then the label changes, then the rest
    std::vector<Curve> curves;
    for (int64_t k = 0; k < g->nfeat;) {
        const int type = (int)g->feat[k];
        arg_check(type >= 1 && type <= 4, "features: unknown curve type");
        const float* q = &g->feat[k + 1];
        const int64_t len = type == 1 ? 4 : type == 2 ? 7 : type == 4 ? 2 + 3 * (int64_t)q[0] : 8;
        arg_check(k + len <= g->nfeat && (type != 4 || q[0] >= 0), "features: a curve beyond the array");
        if (type == 1) {
            fp.insert(fp.end(), { q[0], q[1], q[2] });
        } else if (type != 4 || q[0] >= 2) {
            Curve cv{};
            cv.type = type;
            cv.q = q;
            if (type == 3) {   // an orthonormal frame about the circle's normal
                const float* n = q + 3;
                const float a0 = std::fabs(n[0]) < 0.9f ? 1.0f : 0.0f, a1 = a0 > 0 ? 0.0f : 1.0f;
                cv.u[0] = n[1] * 0 - n[2] * a1;
                cv.u[1] = n[2] * a0 - n[0] * 0;
                cv.u[2] = n[0] * a1 - n[1] * a0;
                const float lu = std::sqrt(cv.u[0] * cv.u[0] + cv.u[1] * cv.u[1] + cv.u[2] * cv.u[2]);
                for (float& x : cv.u) {
                    x /= lu;
                }
                cv.w[0] = n[1] * cv.u[2] - n[2] * cv.u[1];
                cv.w[1] = n[2] * cv.u[0] - n[0] * cv.u[2];
                cv.w[2] = n[0] * cv.u[1] - n[1] * cv.u[0];
            }
            cv.pts.resize(3 * (nsamp + 1));
            for (int j = 0; j <= nsamp; ++j) {
                cv.at((float)j / nsamp, &cv.pts[3 * j]);
            }
            curves.push_back(std::move(cv));
        }
        k += len;
    }
    // every curve's sample masks in one round; then the label changes, each bisected
    // 14 times, all of them a round at a time
    std::vector<float> all;
    for (const Curve& cv : curves) {
        all.insert(all.end(), cv.pts.begin(), cv.pts.end());
    }
    const std::vector<unsigned long long> mk_all = masks(all);
    struct Change {
        int curve;
        unsigned long long before;
        float lo, hi;
    };
    std::vector<std::vector<Change>> changes(curves.size());
    std::vector<Change> flat;
    for (size_t c = 0, o = 0; c < curves.size(); ++c, o += nsamp + 1) {
        for (int j = 1; j <= nsamp; ++j) {
            if (mk_all[o + j] != mk_all[o + j - 1]) {
                flat.push_back({ (int)c, mk_all[o + j - 1], (float)(j - 1) / nsamp, (float)j / nsamp });
            }
        }
    }
    for (int it = 0; it < 14 && !flat.empty(); ++it) {
        std::vector<float> X(3 * flat.size());
        for (size_t t = 0; t < flat.size(); ++t) {
            const float mid = 0.5f * (flat[t].lo + flat[t].hi);
            curves[flat[t].curve].at(mid, &X[3 * t]);
        }
        const std::vector<unsigned long long> mm = masks(X);
        for (size_t t = 0; t < flat.size(); ++t) {
            const float mid = 0.5f * (flat[t].lo + flat[t].hi);
            if (mm[t] == flat[t].before) {
                flat[t].lo = mid;
            } else {
                flat[t].hi = mid;
            }
        }
    }
    for (const Change& ch : flat) {
        changes[ch.curve].push_back(ch);
    }
    for (size_t c = 0; c < curves.size(); ++c) {
        const Curve& cv = curves[c];
        for (const Change& ch : changes[c]) {   // pinned first: an edge crossing an interface
            float x[3];
            cv.at(0.5f * (ch.lo + ch.hi), x);
            tp.insert(tp.end(), x, x + 3);
        }
        // arc length in units of h: n = round(total) segments (>= 1; a circle >= 6)
        std::vector<double> acc(nsamp + 1, 0.0);
        for (int j = 1; j <= nsamp; ++j) {
            const float *a = &cv.pts[3 * (j - 1)], *b = &cv.pts[3 * j];
            const float ds = std::sqrt((b[0] - a[0]) * (b[0] - a[0]) + (b[1] - a[1]) * (b[1] - a[1]) +
                                       (b[2] - a[2]) * (b[2] - a[2]));
            const float m[3] = { 0.5f * (a[0] + b[0]), 0.5f * (a[1] + b[1]), 0.5f * (a[2] + b[2]) };
            acc[j] = acc[j - 1] + ds / std::max(hat(m), 1e-6f);
        }
        const double per = cv.type == 4 ? 2.0 : 1.0;   // (a polyline twice as dense)
        const int ns = std::max(cv.type == 3 ? 6 : 1, (int)std::lround(per * acc[nsamp]));
        for (int sg = cv.type == 3 || cv.type == 4 ? 0 : 1; sg < ns; ++sg) {
            const double target = acc[nsamp] * sg / ns;
            int j = 1;
            while (j < nsamp && acc[j] < target) {
                ++j;
            }
            const double fr = acc[j] > acc[j - 1] ? (target - acc[j - 1]) / (acc[j] - acc[j - 1]) : 0.0;
            for (int a = 0; a < 3; ++a) {
                rp.push_back((float)(cv.pts[3 * (j - 1) + a] + fr * (cv.pts[3 * j + a] - cv.pts[3 * (j - 1) + a])));
            }
        }
    }
    fp.insert(fp.end(), tp.begin(), tp.end());
    fp.insert(fp.end(), rp.begin(), rp.end());
    // the labels round each candidate: the 14 axis and diagonal directions and 128 on a
    // golden spiral, at 0.25 h
    std::vector<float> dirs = { 1, 0, 0, -1, 0, 0, 0, 1, 0, 0, -1, 0, 0, 0, 1, 0, 0, -1,
                                .577f, .577f, .577f, -.577f, .577f, .577f, .577f, -.577f, .577f, .577f, .577f, -.577f,
                                -.577f, -.577f, .577f, -.577f, .577f, -.577f, .577f, -.577f, -.577f, -.577f, -.577f, -.577f };
    for (int i = 0; i < 128; ++i) {
        const double zc = 1 - (2 * i + 1) / 128.0, rr = std::sqrt(std::max(0.0, 1 - zc * zc)), ph = 2.39996322972865 * i;
        dirs.insert(dirs.end(), { (float)(rr * std::cos(ph)), (float)(rr * std::sin(ph)), (float)zc });
    }
    const int nd = (int)(dirs.size() / 3), nc = (int)(fp.size() / 3);
    std::vector<int32_t> labs((size_t)nc * nd);
    if (nc > 0) {
        DevMem tmp;
        tmp.device = mem.device;
        const float* dX = tmp.upload(fp.data(), fp.size());
        const float* dD = tmp.upload(dirs.data(), dirs.size());
        int32_t* dL = tmp.alloc<int32_t>((size_t)nc * nd);
        tnk::k_feat_labels<<<blocks_for((int64_t)nc * nd, tnk::kBlock), tnk::kBlock>>>(f, dX, nc, dD, nd, dL);
        TN_HIP(hipGetLastError());
        TN_HIP(hipMemcpy(labs.data(), dL, sizeof(int32_t) * labs.size(), hipMemcpyDeviceToHost));
    }
    std::vector<std::array<int, 4>> ls(nc);
    std::vector<int> nls(nc, 0);
    const float xmax = (g->nx - 1) * g->vs[0], ymax = (g->ny - 1) * g->vs[1], zmax = (g->nz - 1) * g->vs[2];
    for (int c = 0; c < nc; ++c) {
        const float* x = &fp[3 * (size_t)c];
        if (x[0] < 0 || x[1] < 0 || x[2] < 0 || x[0] > xmax || x[1] > ymax || x[2] > zmax) {
            continue;
        }
        std::map<int, int> cnt;
        for (int d = 0; d < nd; ++d) {
            ++cnt[labs[(size_t)c * nd + d]];
        }
        if (cnt.size() < 2) {
            continue;   // hidden: inside one region
        }
        std::vector<std::pair<int, int>> by(cnt.begin(), cnt.end());
        std::sort(by.begin(), by.end(), [](const std::pair<int, int>& a, const std::pair<int, int>& b) {
            return a.second > b.second || (a.second == b.second && a.first < b.first);
        });
        // (the node's own label a tissue one: 0, the exterior, only as a partner)
        std::stable_partition(by.begin(), by.end(), [](const std::pair<int, int>& a) { return a.first != 0; });
        int m = 0;
        for (const auto& pr : by) {
            if (m < 4) {
                ls[c][m++] = pr.first;
            }
        }
        nls[c] = (m > 0 && ls[c][0] == 0) ? 0 : m;
    }
    // keep one per 0.3 h, in candidate order (corners first)
    std::vector<float> keepP;
    std::vector<std::array<int, 4>> keepL;
    std::vector<int> keepN;
    for (int c = 0; c < nc; ++c) {
        if (nls[c] < 2) {
            continue;
        }
        const float* x = &fp[3 * (size_t)c];
        const float r = 0.3f * hat(x);
        bool near = false;
        for (size_t k = 0; k < keepP.size() / 3 && !near; ++k) {
            const float dx = keepP[3 * k] - x[0], dy = keepP[3 * k + 1] - x[1], dz = keepP[3 * k + 2] - x[2];
            near = dx * dx + dy * dy + dz * dz < r * r;
        }
        if (!near) {
            keepP.insert(keepP.end(), x, x + 3);
            keepL.push_back(ls[c]);
            keepN.push_back(nls[c]);
        }
    }
    const int nk = (int)(keepP.size() / 3);
    if (nk == 0) {
        return s;
    }
    // the seeds that stay, in order, then the feature nodes
    const int n0 = (int)s.n;
    int32_t* keep = ops.counts(n0);
    const float* dK = mem.upload(keepP.data(), keepP.size());
    if (n0 > 0) {
        tnk::k_feat_drop<<<blocks_for(n0, tnk::kBlock), tnk::kBlock>>>(f, n0, s.P, dK, nk, keep);
        TN_HIP(hipGetLastError());
    }
    const int m = ops.scan(keep, n0);
    Seeds o;
    o.n = (int64_t)m + nk;
    o.P = mem.alloc<float>((size_t)3 * o.n);
    o.lab = mem.alloc<uint16_t>((size_t)o.n);
    o.typ = mem.alloc<uint8_t>((size_t)o.n);
    o.part = mem.alloc<uint16_t>((size_t)2 * o.n);
    o.part3 = mem.alloc<uint16_t>((size_t)o.n);
    if (n0 > 0) {
        tnthin::k_compact_nodes<<<blocks_for(n0, tnk::kBlock), tnk::kBlock>>>(keep, n0, s.P, s.lab, s.typ, s.part,
                                                                              s.part3, o.P, o.lab, o.typ, o.part,
                                                                              o.part3);
        TN_HIP(hipGetLastError());
    }
    std::vector<uint16_t> flab(nk), fpart(2 * (size_t)nk), fpart3(nk);
    std::vector<uint8_t> ftyp(nk, TN_CORNER);
    for (int k = 0; k < nk; ++k) {
        const std::array<int, 4>& L = keepL[k];
        const int mm = keepN[k];
        flab[k] = (uint16_t)L[0];
        fpart[2 * k] = (uint16_t)(mm > 1 ? L[1] : TN_NOLAB);
        fpart[2 * k + 1] = (uint16_t)(mm > 2 ? L[2] : TN_NOLAB);
        fpart3[k] = (uint16_t)(mm > 3 ? L[3] : TN_NOLAB);
    }
    TN_HIP(hipMemcpy(o.P + 3 * (size_t)m, keepP.data(), sizeof(float) * keepP.size(), hipMemcpyHostToDevice));
    TN_HIP(hipMemcpy(o.lab + m, flab.data(), sizeof(uint16_t) * nk, hipMemcpyHostToDevice));
    TN_HIP(hipMemcpy(o.typ + m, ftyp.data(), nk, hipMemcpyHostToDevice));
    TN_HIP(hipMemcpy(o.part + 2 * (size_t)m, fpart.data(), sizeof(uint16_t) * 2 * nk, hipMemcpyHostToDevice));
    TN_HIP(hipMemcpy(o.part3 + m, fpart3.data(), sizeof(uint16_t) * nk, hipMemcpyHostToDevice));
    return o;
}

Seeds run_seeding(const tn_grid_view* g, const uint8_t* grade, const tn_relax_params* prm, int dev, DevMem& mem) {
    check_grid(g);
    arg_check(prm != nullptr && grade != nullptr, "seeding: null argument");
    arg_check(g->nfeat >= 0 && (g->nfeat == 0 || g->feat), "grid: feat is null");
    TN_HIP(hipSetDevice(dev));
    mem.device = dev;
    const tnd::Field f = upload_grid(g, mem);
    const uint8_t* d_grade = mem.upload(grade, (size_t)g->nx * g->ny * g->nz);
    const Seeds s = seeding_core(f, d_grade, g->hmin, g->hmax, prm, mem);
    // seed_cpu (src/v2m_particles.cpp:625-635): the shape input's feature nodes last
    return g->nfeat > 0 ? add_feature_nodes(g, f, s, mem) : s;
}

// seed_cpu on a grid already on the device (f, its grades d_grade), scratch in mem
Seeds seeding_core(const tnd::Field& f, const uint8_t* d_grade, float hmin, float hmax, const tn_relax_params* prm,
                   DevMem& mem) {
    arg_check(prm != nullptr, "seeding: null argument");
    arg_check(prm->nseed >= 1 && prm->nseed <= 64, "seeding: nseed out of range");
    arg_check(prm->jseed >= 0.0f && std::isfinite(prm->jseed), "seeding: jseed must be finite and >= 0");
    SeedOps ops{ mem };
    const tnd::Field* g = &f;   // (its dims)
    const int64_t nv = (int64_t)g->nx * g->ny * g->nz;
    // v2m_seed_spacing (src/opencl/v2m_seed_body.cl:445-447), with the host's powf
    std::vector<float> spacing(prm->nseed);
    for (int L = 0; L < prm->nseed; ++L) {
        spacing[L] = prm->nseed <= 1 ? hmin : hmin * std::pow(hmax / hmin, (float)L / (float)(prm->nseed - 1));
        arg_check(spacing[L] > 0.0f && std::isfinite(spacing[L]), "seeding: bad seed spacing");
    }
    (void)nv;
    tnk::SeedArgs sa;
    sa.grade = d_grade;
    sa.spacing = mem.upload(spacing.data(), spacing.size());
    sa.nseed = prm->nseed;
    sa.voxmode = prm->voxel_trap ? 1 : 0;
    tnk::VertexGrid vg;
    vg.vx1 = g->nx + 1;
    vg.vy1 = g->ny + 1;
    vg.nvert = (int64_t)vg.vx1 * vg.vy1 * (g->nz + 1);
    // the corner vertices first: their count sizes the arrays the lattice goes into
    int32_t* clist = nullptr;
    int nc = 0;
    if (prm->corners) {
        int32_t* cflag = ops.counts(vg.nvert);
        tnk::k_vertex_mark<1><<<blocks_for(vg.nvert, tnk::kBlock), tnk::kBlock>>>(f, vg, cflag);
        nc = ops.scan(cflag, vg.nvert);
        if (nc > 0) {
            clist = mem.alloc<int32_t>(nc);
            tnk::k_compact_index<<<blocks_for(vg.nvert, tnk::kBlock), tnk::kBlock>>>(cflag, vg.nvert, clist);
        }
    }
    int32_t* offs = ops.counts(nv);
    tnk::k_seed_count<<<blocks_for(nv, tnk::kBlock), tnk::kBlock>>>(f, sa, offs);
    const int32_t total = ops.scan(offs, nv);
    const bool junctions = prm->jseed > 0.0f;
    Seeds lat;
    lat.n = total;
    const size_t cap = (size_t)std::max<int64_t>(junctions ? total : (int64_t)total + nc, 1);
    lat.P = mem.alloc<float>(3 * cap);
    lat.lab = mem.alloc<uint16_t>(cap);
    lat.typ = mem.alloc<uint8_t>(cap);
    lat.part = mem.alloc<uint16_t>(2 * cap);
    lat.part3 = mem.alloc<uint16_t>(cap);
    if (total > 0) {
        tnk::k_seed_fill<<<blocks_for(nv, tnk::kBlock), tnk::kBlock>>>(f, sa, offs, lat.P, lat.lab);
        tnk::k_seed_classify<<<blocks_for(total, tnk::kBlock), tnk::kBlock>>>(f, total, lat.P, lat.lab, lat.typ,
                                                                              lat.part);
        TN_HIP(hipMemset(lat.part3, 0xFF, sizeof(uint16_t) * total));   // kNoLab
        TN_HIP(hipGetLastError());
    }
    Seeds s = lat;
    if (junctions) {
        junction_pass(hmax, f, sa, prm->jseed, vg, lat, nc, ops, s);
    }
    if (nc > 0) {
        tnk::SeedNodes o{ s.P, s.lab, s.typ, s.part, s.part3 };
        tnk::k_corner_fill<<<blocks_for(nc, tnk::kBlock), tnk::kBlock>>>(f, vg, clist, nc, o, (int)s.n);
        TN_HIP(hipGetLastError());
        s.n += nc;
    }
    TN_HIP(hipDeviceSynchronize());
    return s;
}

// Every plan's iterations as grouped launches on one stream: each of an iteration's
// kernels is launched once for all the plans still running (a device table holds
// their arguments), one copy brings all their control blocks to the host, and the
// plans that asked for a rebuild get it on the same stream before the next
// iteration. Each plan computes exactly what its own launches would.
template <class Plans>
void run_grouped(Plans& plans, int64_t& launches, int64_t& steps) {
    const int count = (int)plans.size();
    hipStream_t bs = plans[0]->stream;
    bool all_plain = true;   // every grid a label volume
    for (int k = 0; k < count; ++k) {
        all_plain = all_plain && plans[k]->f.gm == 0;
    }
    std::vector<hipStream_t> own(count);
    for (int k = 0; k < count; ++k) {
        own[k] = plans[k]->stream;
    }
    struct RestoreStreams {
        Plans& p;
        std::vector<hipStream_t>& s;
        ~RestoreStreams() {
            for (size_t k = 0; k < s.size(); ++k) {
                p[k]->stream = s[k];
            }
        }
    } restore{ plans, own };
    TN_HIP(hipDeviceSynchronize());   // every plan's set-up work is done
    for (int k = 0; k < count; ++k) {
        plans[k]->stream = bs;
        plans[k]->overlap_moves = false;
    }
    DevMem gm;
    gm.device = plans[0]->device;
    tnk::PlanIter* dtab = gm.alloc<tnk::PlanIter>(count);
    tnk::Ctl* dctl = gm.alloc<tnk::Ctl>(count);
    // pinned host copies: the table rows, the control blocks, the block maps
    struct Pinned {
        void* p = nullptr;
        ~Pinned() {
            if (p) {
                (void)hipHostFree(p);
            }
        }
    };
    Pinned ptab, pctl, pmap;
    TN_HIP(hipHostMalloc(&ptab.p, sizeof(tnk::PlanIter) * count, hipHostMallocDefault));
    TN_HIP(hipHostMalloc(&pctl.p, sizeof(tnk::Ctl) * count, hipHostMallocDefault));
    auto* htab = static_cast<tnk::PlanIter*>(ptab.p);
    auto* hctl = static_cast<tnk::Ctl*>(pctl.p);
    for (int k = 0; k < count; ++k) {
        htab[k] = plans[k]->iter_args();
    }
    TN_HIP(hipMemcpyAsync(dtab, htab, sizeof(tnk::PlanIter) * count, hipMemcpyHostToDevice, bs));
    // block maps: for each grid shape, the plan of every block and each plan's first block
    int64_t map_total = 0;
    for (int k = 0; k < count; ++k) {
        for (int sh = 0; sh < tnk::kShapes; ++sh) {
            map_total += htab[k].nblk[sh];
        }
    }
    map_total += (int64_t)tnk::kShapes * count;
    TN_HIP(hipHostMalloc(&pmap.p, sizeof(int32_t) * map_total, hipHostMallocDefault));
    int32_t* hmap = static_cast<int32_t*>(pmap.p);
    int32_t* dmap = gm.alloc<int32_t>((size_t)map_total);
    const int32_t* bplan[tnk::kShapes];
    const int32_t* bfirst[tnk::kShapes];
    unsigned nblocks[tnk::kShapes];
    std::vector<int> active;
    for (int k = 0; k < count; ++k) {
        if (plans[k]->more()) {
            active.push_back(k);
        }
    }
    auto build_maps = [&]() {
        int64_t o = 0;
        for (int sh = 0; sh < tnk::kShapes; ++sh) {
            int32_t* first = hmap + o;
            o += count;
            int32_t* plan = hmap + o;
            int nb = 0;
            for (int k : active) {
                first[k] = nb;
                for (int b = 0; b < htab[k].nblk[sh]; ++b) {
                    plan[nb++] = k;
                }
            }
            bfirst[sh] = dmap + (first - hmap);
            bplan[sh] = dmap + (plan - hmap);
            nblocks[sh] = (unsigned)nb;
            o += nb;
        }
        TN_HIP(hipMemcpyAsync(dmap, hmap, sizeof(int32_t) * o, hipMemcpyHostToDevice, bs));
    };
    build_maps();
    // Launch-shape tuning: the general move passes' nodes per wave. The candidates
    // take turns (round robin) on grouped iterations 1 .. ncand * kTuneIters (the
    // first iteration, cold, is skipped), each timed by events around the two general
    // passes; the one with the least time per queued general-pass node (the plans'
    // nmove_rest + nmove_def, which shrink as nodes settle) is kept. The nodes'
    // results do not depend on it.
    static constexpr int kTuneLanes[] = { 1, 2, 4, 8, 16 };
    constexpr int kTuneIters = 2;
    const int ncand = (int)(sizeof kTuneLanes / sizeof kTuneLanes[0]);
    std::vector<double> tune_ms(ncand, 0.0), tune_nodes(ncand, 0.0);
    hipEvent_t tune_ev[4];
    for (hipEvent_t& e : tune_ev) {
        TN_HIP(hipEventCreate(&e));
    }
    struct EventsGone {
        hipEvent_t* e;
        ~EventsGone() {
            for (int k = 0; k < 4; ++k) {
                (void)hipEventDestroy(e[k]);
            }
        }
    } events_gone{ tune_ev };
    // the side stream of the junction queues' general pass, and its fork and join
    hipStream_t gs = plans[0]->side[0];
    hipEvent_t fork_ev = plans[0]->ev_force, join_ev = plans[0]->ev_rest;
    int tune_step = std::getenv("TN_NO_TUNE") ? ncand * kTuneIters : -1;   // -1: the untimed first iteration
    auto set_lanes = [&](int lanes) {
        for (int k = 0; k < count; ++k) {
            plans[k]->rest_lanes = lanes;
            htab[k].ma.lanes = lanes;
        }
        TN_HIP(hipMemcpyAsync(dtab, htab, sizeof(tnk::PlanIter) * count, hipMemcpyHostToDevice, bs));
    };
    if (tune_step < ncand * kTuneIters) {
        set_lanes(kTuneLanes[0]);
    }
    while (!active.empty()) {
        using namespace tnk;
        const bool tuning = tune_step >= 0 && tune_step < ncand * kTuneIters;
        kg_iter_begin<<<nblocks[kShOne], 64, 0, bs>>>(dtab, bplan[kShOne], bfirst[kShOne]);
        kg_force<<<nblocks[kShNodes], kBlock, 0, bs>>>(dtab, bplan[kShNodes], bfirst[kShNodes]);
        // The general pass over every volume's junction queue runs on a side stream from
        // the forces on, beside the easy and lean passes (they move disjoint nodes, as in
        // a single plan's iteration); the pass over the handed-back nodes follows them.
        auto moves = [&](auto plain) {
            constexpr bool kP = decltype(plain)::value;
            TN_HIP(hipEventRecord(fork_ev, bs));
            TN_HIP(hipStreamWaitEvent(gs, fork_ev, 0));
            if (tuning) {
                TN_HIP(hipEventRecord(tune_ev[0], gs));
            }
            kg_move<0, kP><<<nblocks[kShGen], kBlock, 0, gs>>>(dtab, bplan[kShGen], bfirst[kShGen]);
            if (tuning) {
                TN_HIP(hipEventRecord(tune_ev[1], gs));
            }
            TN_HIP(hipEventRecord(join_ev, gs));
            kg_move_easy<kP><<<nblocks[kShNodes], kBlock, 0, bs>>>(dtab, bplan[kShNodes], bfirst[kShNodes]);
            kg_move<1, kP><<<nblocks[kShLean], kBlock, 0, bs>>>(dtab, bplan[kShLean], bfirst[kShLean]);
            kg_move<2, kP><<<nblocks[kShLean], kBlock, 0, bs>>>(dtab, bplan[kShLean], bfirst[kShLean]);
            if (tuning) {
                TN_HIP(hipEventRecord(tune_ev[2], bs));
            }
            kg_move<3, kP><<<nblocks[kShGen], kBlock, 0, bs>>>(dtab, bplan[kShGen], bfirst[kShGen]);
            if (tuning) {
                TN_HIP(hipEventRecord(tune_ev[3], bs));
            }
            TN_HIP(hipStreamWaitEvent(bs, join_ev, 0));
        };
        if (all_plain) {
            moves(std::true_type{});
        } else {
            moves(std::false_type{});
        }
        kg_stats<<<nblocks[kShParts], kBlock, 0, bs>>>(dtab, bplan[kShParts], bfirst[kShParts]);
        kg_control<<<nblocks[kShOne], kBlock, 0, bs>>>(dtab, bplan[kShOne], bfirst[kShOne]);
        kg_refresh_list<<<nblocks[kShNodes], kBlock, 0, bs>>>(dtab, bplan[kShNodes], bfirst[kShNodes]);
        kg_truss_node<<<nblocks[kShRefresh], kWave, 0, bs>>>(dtab, bplan[kShRefresh], bfirst[kShRefresh]);
        kg_ctl_gather<<<nblocks[kShOne], 64, 0, bs>>>(dtab, bplan[kShOne], bfirst[kShOne], dctl);
        TN_HIP(hipGetLastError());
        TN_HIP(hipMemcpyAsync(hctl, dctl, sizeof(tnk::Ctl) * count, hipMemcpyDeviceToHost, bs));
        TN_HIP(hipStreamSynchronize(bs));
        ++launches;
        if (tuning) {
            float a = 0.0f, b = 0.0f;
            TN_HIP(hipEventElapsedTime(&a, tune_ev[0], tune_ev[1]));
            TN_HIP(hipEventElapsedTime(&b, tune_ev[2], tune_ev[3]));
            double queued = 0.0;
            for (const int k : active) {
                queued += (double)hctl[k].nmove_rest + (double)hctl[k].nmove_def;
            }
            tune_ms[tune_step % ncand] += (double)a + (double)b;
            tune_nodes[tune_step % ncand] += queued;
        }
        if (tune_step < ncand * kTuneIters) {
            ++tune_step;
            if (tune_step == ncand * kTuneIters) {   // keep the least time per queued node
                auto per_node = [&](int c) { return tune_ms[c] / (tune_nodes[c] > 1.0 ? tune_nodes[c] : 1.0); };
                int best = 0;
                for (int c = 1; c < ncand; ++c) {
                    best = per_node(c) < per_node(best) ? c : best;
                }
                set_lanes(kTuneLanes[best]);
                if (std::getenv("TN_BATCH_VERBOSE")) {
                    std::fprintf(stderr, "[tn_relax_batch] general pass lanes (us per queued node):");
                    for (int c = 0; c < ncand; ++c) {
                        std::fprintf(stderr, " %d: %.3f (%.3f ms, %.0f nodes)", kTuneLanes[c], 1e3 * per_node(c),
                                     tune_ms[c] / kTuneIters, tune_nodes[c] / kTuneIters);
                    }
                    std::fprintf(stderr, " -> %d\n", kTuneLanes[best]);
                }
            } else {
                set_lanes(kTuneLanes[tune_step % ncand]);
            }
        }
        bool changed = false;
        for (size_t a = 0; a < active.size();) {
            const int k = active[a];
            tn_plan* p = &*plans[k];
            *p->hCtl = hctl[k];
            ++steps;
            const int before = p->rebuilds;
            // a rebuild runs on the plan's own stream (bs is idle here), so the plans'
            // rebuilds overlap; the next grouped launch waits for it
            p->stream = own[k];
            const bool go = p->finish_iteration() && p->more();
            p->stream = bs;
            if (p->rebuilds != before) {   // the rebuild swapped the node buffers
                TN_HIP(hipEventRecord(p->ev_iter, own[k]));
                TN_HIP(hipStreamWaitEvent(bs, p->ev_iter, 0));
                htab[k] = p->iter_args();
                TN_HIP(hipMemcpyAsync(dtab + k, htab + k, sizeof(tnk::PlanIter), hipMemcpyHostToDevice, bs));
            }
            if (go) {
                ++a;
            } else {
                active[a] = active.back();
                active.pop_back();
                changed = true;
            }
        }
        if (changed && !active.empty()) {
            TN_HIP(hipStreamSynchronize(bs));   // the maps' host copy is about to change
            std::sort(active.begin(), active.end());
            build_maps();
        }
    }
    TN_HIP(hipStreamSynchronize(bs));
}

// ---- node thinning and feature protection (tn_thin.hpp) ------------------------------

// the h field of a grid view as a Field (v2m_h_at reads only h and the voxel sizes)
tnd::Field size_field(const tn_grid_view* g, DevMem& mem) {
    tnd::Field f{};
    f.nx = g->nx;
    f.ny = g->ny;
    f.nz = g->nz;
    f.nbx = g->nbx;
    f.nby = g->nby;
    f.nbz = g->nbz;
    f.vx = g->vs[0];
    f.vy = g->vs[1];
    f.vz = g->vs[2];
    set_reciprocals(f);
    f.h = mem.upload(g->h, (size_t)g->nx * g->ny * g->nz);
    return f;
}

// v2m_h_at on the host, src/opencl/v2m_particle_body.cl:30-36
float host_h_at(const tn_grid_view* g, float px, float py, float pz) {
    int i = (int)std::floor(px / g->vs[0] + 0.5f), j = (int)std::floor(py / g->vs[1] + 0.5f),
        k = (int)std::floor(pz / g->vs[2] + 0.5f);
    i = i < 0 ? 0 : (i >= g->nx ? g->nx - 1 : i);
    j = j < 0 ? 0 : (j >= g->ny ? g->ny - 1 : j);
    k = k < 0 ? 0 : (k >= g->nz ? g->nz - 1 : k);
    return g->h[i + (size_t)g->nx * (j + (size_t)g->ny * k)];
}

// The nodes of a compaction on the device (null stream), and the kept ones back into
// the host arrays, in order. keep[0..n) holds the flags; returns the kept count.
struct NodeArraysDev {
    float* P = nullptr;
    uint16_t *lab = nullptr, *part = nullptr, *part3 = nullptr;
    uint8_t* typ = nullptr;
};

NodeArraysDev upload_nodes(int n, const float* P, const uint16_t* lab, const uint8_t* typ, const uint16_t* part,
                           const uint16_t* part3, DevMem& mem) {
    NodeArraysDev d;
    d.P = mem.upload(P, (size_t)3 * n);
    d.lab = mem.upload(lab, (size_t)n);
    d.typ = mem.upload(typ, (size_t)n);
    d.part = mem.upload(part, (size_t)2 * n);
    if (part3) {
        d.part3 = mem.upload(part3, (size_t)n);
    } else {   // no third partners
        const std::vector<uint16_t> none((size_t)n, (uint16_t)TN_NOLAB);
        d.part3 = mem.upload(none.data(), (size_t)n);
    }
    return d;
}

int64_t compact_back(const NodeArraysDev& d, int n, int32_t* keep, SeedOps& ops, float* P, uint16_t* lab,
                     uint8_t* typ, uint16_t* part, uint16_t* part3) {
    DevMem& mem = ops.mem;
    const int m = ops.scan(keep, n);
    if (m == n) {
        return n;
    }
    NodeArraysDev o;
    o.P = mem.alloc<float>((size_t)3 * std::max(m, 1));
    o.lab = mem.alloc<uint16_t>(std::max(m, 1));
    o.typ = mem.alloc<uint8_t>(std::max(m, 1));
    o.part = mem.alloc<uint16_t>((size_t)2 * std::max(m, 1));
    o.part3 = mem.alloc<uint16_t>(std::max(m, 1));
    tnthin::k_compact_nodes<<<blocks_for(n, tnk::kBlock), tnk::kBlock>>>(keep, n, d.P, d.lab, d.typ, d.part, d.part3, o.P,
                                                                         o.lab, o.typ, o.part, o.part3);
    TN_HIP(hipGetLastError());
    if (m > 0) {
        TN_HIP(hipMemcpy(P, o.P, sizeof(float) * 3 * m, hipMemcpyDeviceToHost));
        TN_HIP(hipMemcpy(lab, o.lab, sizeof(uint16_t) * m, hipMemcpyDeviceToHost));
        TN_HIP(hipMemcpy(typ, o.typ, m, hipMemcpyDeviceToHost));
        TN_HIP(hipMemcpy(part, o.part, sizeof(uint16_t) * 2 * m, hipMemcpyDeviceToHost));
        if (part3) {
            TN_HIP(hipMemcpy(part3, o.part3, sizeof(uint16_t) * m, hipMemcpyDeviceToHost));
        }
    }
    return m;
}

// thin_nodes (src/v2m_particles.cpp:1223-1357) with factor B on device dev: the host
// arrays are compacted in place; returns the kept count
int64_t thin_on_device(const tn_grid_view* g, float B, int n, float* P, uint16_t* lab, uint8_t* typ, uint16_t* part,
                       uint16_t* part3, int dev) {
    TN_HIP(hipSetDevice(dev));
    DevMem mem;
    mem.device = dev;
    SeedOps ops{ mem };
    const tnd::Field f = size_field(g, mem);
    const NodeArraysDev d = upload_nodes(n, P, lab, typ, part, part3, mem);
    float4* X = mem.alloc<float4>(n);
    int32_t* cand = ops.counts(n);
    int32_t* hkey = mem.alloc<int32_t>(n);
    int32_t* state = mem.alloc<int32_t>(n);
    const unsigned nb = blocks_for(n, tnk::kBlock);
    tnthin::k_thin_init<<<nb, tnk::kBlock>>>(f, d.P, d.typ, n, X, cand, hkey, state);
    TN_HIP(hipGetLastError());
    const int m = ops.scan(cand, n);
    int32_t* rank = mem.alloc<int32_t>(n);
    tnk::k_fill_i32<<<nb, tnk::kBlock>>>(rank, n, -1);
    int32_t* order = nullptr;   // the candidates in the greedy order
    if (m > 0) {
        // interfaces first, by h, by index: stable passes over h, then interior-ness
        int32_t* list = mem.alloc<int32_t>(m);
        tnk::k_compact_index<<<nb, tnk::kBlock>>>(cand, n, list);
        int32_t* keys = mem.alloc<int32_t>(m);
        const unsigned mb = blocks_for(m, tnk::kBlock);
        tnthin::k_thin_gather_keys<<<mb, tnk::kBlock>>>(hkey, list, m, keys);
        auto s1 = ops.sort(keys, list, m, 32);
        tnthin::k_thin_interior<<<nb, tnk::kBlock>>>(d.typ, n, hkey);   // hkey now: interior flags
        int32_t* keys2 = mem.alloc<int32_t>(m);
        tnthin::k_thin_gather_keys<<<mb, tnk::kBlock>>>(hkey, s1.second, m, keys2);
        order = ops.sort(keys2, s1.second, m, 1).second;
        tnthin::k_thin_rank<<<mb, tnk::kBlock>>>(order, m, rank);
        TN_HIP(hipGetLastError());
    }
    // the hash of src/v2m_particles.cpp:1240-1255, and a superset grid per level that a
    // candidate can take: nodes of h in [min h, max h] of the grid
    tnthin::Levels Lv{};
    Lv.B = B;
    Lv.c0 = std::max(1e-6f, B * g->hmin);
    Lv.nlev = 1;
    while (Lv.nlev < 12 && Lv.c0 * (float)(1 << (Lv.nlev - 1)) < B * g->hmax) {
        ++Lv.nlev;
    }
    float lo[3] = { INFINITY, INFINITY, INFINITY }, hi[3] = { -INFINITY, -INFINITY, -INFINITY };
    for (int i = 0; i < n; ++i) {
        for (int a = 0; a < 3; ++a) {
            lo[a] = std::min(lo[a], P[3 * i + a]);
            hi[a] = std::max(hi[a], P[3 * i + a]);
        }
    }
    float ext = 0.0f;
    for (int a = 0; a < 3; ++a) {
        Lv.lo[a] = lo[a];
        ext = std::max(ext, hi[a] - lo[a]);
    }
    const size_t nv = (size_t)g->nx * g->ny * g->nz;
    float hlo = INFINITY, hhi = 0.0f;
    for (size_t v = 0; v < nv; ++v) {
        hlo = std::min(hlo, g->h[v]);
        hhi = std::max(hhi, g->h[v]);
    }
    auto level_of = [&](float r) {
        int L = 0;
        while (L + 1 < Lv.nlev && Lv.c0 * (float)(1 << L) < r) {
            ++L;
        }
        return L;
    };
    const int Lmin = level_of(B * hlo), Lmax = level_of(B * hhi);
    int64_t ncells = 0;
    for (int L = 0; L < Lv.nlev; ++L) {
        const float c = Lv.c0 * (float)(1 << L);
        Lv.C[L] = std::max(c, ext / 256.0f);
        Lv.off[L] = ncells;
        for (int a = 0; a < 3; ++a) {
            Lv.dim[L][a] = L < Lmin || L > Lmax ? 1 : (int)std::floor((hi[a] - lo[a]) / Lv.C[L]) + 1;
        }
        if (L >= Lmin && L <= Lmax) {
            ncells += (int64_t)Lv.dim[L][0] * Lv.dim[L][1] * Lv.dim[L][2];
        }
    }
    arg_check(ncells < ((int64_t)1 << 31) - 1, "thin: too many cells");
    int32_t* cstart = ops.counts(ncells);
    int32_t* clist = mem.alloc<int32_t>((size_t)n * (Lmax - Lmin + 1));
    int32_t* fill = ops.counts(ncells);
    for (int L = Lmin; L <= Lmax; ++L) {
        tnthin::k_thin_count_cells<<<nb, tnk::kBlock>>>(Lv, X, n, L, cstart);
    }
    ops.scan(cstart, ncells);
    for (int L = Lmin; L <= Lmax; ++L) {
        tnthin::k_thin_fill_cells<<<nb, tnk::kBlock>>>(Lv, X, n, L, cstart, fill, clist);
    }
    TN_HIP(hipGetLastError());
    if (m > 0) {   // each candidate's earlier conflicts, then the rounds
        tnthin::Nodes tn{ X, d.lab, d.typ, d.part, d.part3, rank, n };
        const unsigned mb = blocks_for(m, tnk::kBlock);
        int32_t* coff = ops.counts(m);
        tnthin::k_thin_count<<<mb, tnk::kBlock>>>(Lv, tn, cstart, clist, order, m, coff);
        const int ncon = ops.scan(coff, m);
        int32_t* con = mem.alloc<int32_t>(std::max(ncon, 1));
        tnthin::k_thin_fill<<<mb, tnk::kBlock>>>(Lv, tn, cstart, clist, order, m, coff, con);
        int32_t* left = mem.alloc<int32_t>(1);
        int32_t open = m;
        for (int round = 0; open > 0; ++round) {
            if (round > m) {
                throw TnError(TN_ERR_INTERNAL, "thin: the rounds do not converge");
            }
            tnk::k_fill_i32<<<1, tnk::kBlock>>>(left, 1, 0);
            tnthin::k_thin_round<<<mb, tnk::kBlock>>>(order, m, coff, con, state, left);
            TN_HIP(hipGetLastError());
            TN_HIP(hipMemcpy(&open, left, sizeof open, hipMemcpyDeviceToHost));
        }
    }
    int32_t* keep = ops.counts(n);
    tnthin::k_thin_keep<<<nb, tnk::kBlock>>>(state, n, keep);
    return compact_back(d, n, keep, ops, P, lab, typ, part, part3);
}

// protect_features (src/v2m_particles.cpp:1002-1221) with the feature curves feat on
// device dev. The balls between neighbouring pinned nodes on each curve are built on
// the host as the reference builds them (a few per curve); the test of every node
// against the balls of its cell, and the compaction, run on the device.
int64_t protect_on_device(const tn_grid_view* g, const float* feat, int64_t nfeat, int n, float* P, uint16_t* lab,
                          uint8_t* typ, uint16_t* part, uint16_t* part3, int dev) {
    std::vector<int> pins;
    for (int i = 0; i < n; ++i) {
        if (typ[i] == TN_CORNER) {
            pins.push_back(i);
        }
    }
    // locate (src/v2m_particles.cpp:1018-1084): distance of p from curve `type` q, and
    // its parameter along it
    auto seg = [](const float* a, const float* b, const float* p, float& dd, float& tt) {
        float ab[3], ap[3], ab2 = 0, apab = 0;
        for (int k = 0; k < 3; ++k) {
            ab[k] = b[k] - a[k];
            ap[k] = p[k] - a[k];
            ab2 += ab[k] * ab[k];
            apab += ap[k] * ab[k];
        }
        tt = ab2 > 0 ? std::min(1.0f, std::max(0.0f, apab / ab2)) : 0.0f;
        dd = 0;
        for (int k = 0; k < 3; ++k) {
            const float e = ap[k] - tt * ab[k];
            dd += e * e;
        }
        dd = std::sqrt(dd);
    };
    auto locate = [&](int type, const float* q, const float* p, float& dist, float& t) {
        if (type == 2) {
            seg(q, q + 3, p, dist, t);
        } else if (type == 4) {
            const int m = (int)q[0];
            dist = 1e30f;
            for (int i = 0; i + 1 < m; ++i) {
                float dd, tt;
                seg(q + 1 + 3 * i, q + 4 + 3 * i, p, dd, tt);
                if (dd < dist) {
                    dist = dd;
                    t = (float)i + tt;
                }
            }
        } else {
            const float *c = q, *nn = q + 3;
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
            rho = std::sqrt(rho);
            dist = std::hypot(h, rho - q[6]);
            const float a0 = std::fabs(nn[0]) < 0.9f ? 1.0f : 0.0f, a1 = a0 > 0 ? 0.0f : 1.0f;
            float u[3] = { -nn[2] * a1, nn[2] * a0, nn[0] * a1 - nn[1] * a0 };
            const float lu = std::sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
            for (float& x : u) {
                x /= lu;
            }
            const float w[3] = { nn[1] * u[2] - nn[2] * u[1], nn[2] * u[0] - nn[0] * u[2], nn[0] * u[1] - nn[1] * u[0] };
            t = std::atan2(rad[0] * w[0] + rad[1] * w[1] + rad[2] * w[2], rad[0] * u[0] + rad[1] * u[1] + rad[2] * u[2]);
        }
    };
    std::vector<float4> balls;   // centre, radius^2
    for (int64_t k = 0; k < nfeat;) {
        const int type = (int)feat[k];
        arg_check(type >= 1 && type <= 4, "features: unknown curve type");
        const float* q = &feat[k + 1];
        const int64_t len = type == 1 ? 4 : type == 2 ? 7 : type == 4 ? 2 + 3 * (int64_t)q[0] : 8;
        arg_check(k + len <= nfeat && (type != 4 || q[0] >= 0), "features: a curve beyond the array");
        k += len;
        if (type == 1 || (type == 4 && q[0] < 2)) {
            continue;
        }
        std::vector<std::pair<float, int>> on;
        for (int j : pins) {
            const float* p = &P[3 * (size_t)j];
            float dist, t = 0;
            locate(type, q, p, dist, t);
            if (dist < 1e-3f * host_h_at(g, p[0], p[1], p[2])) {
                on.emplace_back(t, j);
            }
        }
        std::sort(on.begin(), on.end());
        const size_t m = on.size();
        for (size_t i = 0; i + 1 < m + (type == 3 && m > 2 ? 1 : 0); ++i) {   // (a circle closes)
            const float *a = &P[3 * (size_t)on[i].second], *b = &P[3 * (size_t)on[(i + 1) % m].second];
            const float c[3] = { 0.5f * (a[0] + b[0]), 0.5f * (a[1] + b[1]), 0.5f * (a[2] + b[2]) };
            const float r2 = 0.25f * ((a[0] - b[0]) * (a[0] - b[0]) + (a[1] - b[1]) * (a[1] - b[1]) +
                                      (a[2] - b[2]) * (a[2] - b[2]));
            const float h = host_h_at(g, c[0], c[1], c[2]);
            if (r2 > 0 && r2 <= 0.64f * h * h) {
                balls.push_back(make_float4(c[0], c[1], c[2], r2));
            }
        }
    }
    if (balls.empty()) {
        return n;
    }
    // the balls by cells of the largest diameter (src/v2m_particles.cpp:1122-1157)
    float cs = 0;
    for (const float4& bl : balls) {
        cs = std::max(cs, 2.0f * std::sqrt(bl.w));
    }
    auto key = [](int64_t ix, int64_t iy, int64_t iz) {
        return (uint64_t)(ix + (1 << 20)) << 42 | (uint64_t)(iy + (1 << 20)) << 21 | (uint64_t)(iz + (1 << 20));
    };
    std::vector<std::pair<uint64_t, int32_t>> cells;
    for (size_t bi = 0; bi < balls.size(); ++bi) {
        const float bl[4] = { balls[bi].x, balls[bi].y, balls[bi].z, balls[bi].w };
        const float r = std::sqrt(bl[3]);
        int64_t clo[3], chi[3];
        for (int k = 0; k < 3; ++k) {
            clo[k] = (int64_t)std::floor((bl[k] - r) / cs);
            chi[k] = (int64_t)std::floor((bl[k] + r) / cs);
        }
        for (int64_t x = clo[0]; x <= chi[0]; ++x) {
            for (int64_t y = clo[1]; y <= chi[1]; ++y) {
                for (int64_t z = clo[2]; z <= chi[2]; ++z) {
                    cells.emplace_back(key(x, y, z), (int32_t)bi);
                }
            }
        }
    }
    std::sort(cells.begin(), cells.end());
    std::vector<uint64_t> ckey(cells.size());
    std::vector<int32_t> cball(cells.size());
    for (size_t e = 0; e < cells.size(); ++e) {
        ckey[e] = cells[e].first;
        cball[e] = cells[e].second;
    }
    TN_HIP(hipSetDevice(dev));
    DevMem mem;
    mem.device = dev;
    SeedOps ops{ mem };
    const NodeArraysDev d = upload_nodes(n, P, lab, typ, part, part3, mem);
    const uint64_t* dkey = mem.upload(ckey.data(), ckey.size());
    const int32_t* dball = mem.upload(cball.data(), cball.size());
    const float4* dballs = mem.upload(balls.data(), balls.size());
    int32_t* keep = ops.counts(n);
    tnthin::k_protect<<<blocks_for(n, tnk::kBlock), tnk::kBlock>>>(d.P, d.typ, n, cs, dkey, dball, (int)ckey.size(),
                                                                   dballs, keep);
    TN_HIP(hipGetLastError());
    return compact_back(d, n, keep, ops, P, lab, typ, part, part3);
}

}  // namespace

// The device memory of tn_relax_noalloc: one array per GPU, held from the library's
// load until the process ends. Host code never passes `scratch` itself to a kernel
// or a copy (on the host it only names the symbol); it uses the address that
// hipGetSymbolAddress gives on each device (scratch_base).
static __device__ __align__(256) unsigned char scratch[TN_NOALLOC_SCRATCH_BYTES];

namespace {

// Per device: the array's address there (looked up once, with the device current)
// and the lock that serialises tn_relax_noalloc calls on it (they share the array).
struct ScratchSlot {
    std::mutex lock;
    void* base = nullptr;
};

ScratchSlot& scratch_slot(int dev) {
    static std::mutex slots_lock;
    static std::vector<std::unique_ptr<ScratchSlot>> slots;
    std::lock_guard<std::mutex> g(slots_lock);
    if ((int)slots.size() <= dev) {
        slots.resize(dev + 1);
    }
    if (!slots[dev]) {
        slots[dev].reset(new ScratchSlot);
    }
    return *slots[dev];
}

// the array's address on device dev (current); the caller holds the slot's lock
void* scratch_base(ScratchSlot& s) {
    if (!s.base) {
        void* p = nullptr;
        TN_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) {
            throw TnError(TN_ERR_HIP, "hipGetSymbolAddress(scratch): null address");
        }
        s.base = p;
    }
    return s.base;
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
        int64_t nkeys = 0;
        const tnd::Hash H = make_hash(g, prm, &nkeys);
        *out = plan_build(prm, seeds, dev, H, nkeys, [&](DevMem& mem) { return upload_grid(g, mem); });
    });
}

int tn_plan_set_timing(tn_plan* p, int32_t on) {
    return tn_plan_set_timing_mask(p, on != 0 ? kAllClasses : 0u);
}

int tn_plan_set_timing_mask(tn_plan* p, uint32_t mask) {
    return guarded([&] {
        arg_check(p != nullptr, "null plan");
        TN_HIP(hipSetDevice(p->device));
        TN_HIP(hipStreamSynchronize(p->stream));
        p->timer.collect();
        p->timer.clear();
        p->timer.mask = mask & kAllClasses;
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
                                                                               p->prm.fscale, p->prm.fsurf, n, p->dF,
                                                                               p->dCtl, nullptr, nullptr, nullptr);
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

int tn_plan_field_points(tn_plan* p, int32_t a, int32_t b, int64_t m, const float* X, float* phi, float* psi,
                         float* grad, int32_t* label, int32_t* second) {
    return guarded([&] {
        arg_check(p != nullptr && m >= 0 && m < (1ll << 28) && (m == 0 || X != nullptr), "null argument or bad count");
        arg_check(a >= 0 && a < TN_NOLAB && b >= 0 && b < TN_NOLAB, "label out of range");
        if (m == 0) {
            return;
        }
        TN_HIP(hipSetDevice(p->device));
        DevMem mem;
        mem.device = p->device;
        float* dX = mem.upload(X, (size_t)3 * m);
        float* dphi = phi ? mem.alloc<float>(m) : nullptr;
        float* dpsi = psi ? mem.alloc<float>(m) : nullptr;
        float* dgrad = grad ? mem.alloc<float>((size_t)3 * m) : nullptr;
        int32_t* dlab = label ? mem.alloc<int32_t>(m) : nullptr;
        int32_t* dsec = second ? mem.alloc<int32_t>(m) : nullptr;
        tnk::k_field_points<<<blocks_for(m, tnk::kBlock), tnk::kBlock, 0, p->stream>>>(p->f, a, b, (int)m, dX, dphi,
                                                                                       dpsi, dgrad, dlab, dsec);
        TN_HIP(hipGetLastError());
        TN_HIP(hipStreamSynchronize(p->stream));
        auto down = [&](void* h, const void* d, size_t bytes) {
            if (h) {
                TN_HIP(hipMemcpy(h, d, bytes, hipMemcpyDeviceToHost));
            }
        };
        down(phi, dphi, sizeof(float) * m);
        down(psi, dpsi, sizeof(float) * m);
        down(grad, dgrad, sizeof(float) * 3 * m);
        down(label, dlab, sizeof(int32_t) * m);
        down(second, dsec, sizeof(int32_t) * m);
    });
}

int tn_seed_count(const tn_grid_view* g, const uint8_t* grade, const tn_relax_params* prm, int32_t device,
                  int64_t* n) {
    return guarded([&] {
        arg_check(n != nullptr, "null count");
        DevMem mem;
        *n = run_seeding(g, grade, prm, pick_device(device), mem).n;
    });
}

int tn_seed(const tn_grid_view* g, const uint8_t* grade, const tn_relax_params* prm, int32_t device, int64_t n,
            float* P, uint16_t* lab, uint8_t* typ, uint16_t* part, uint16_t* part3) {
    return guarded([&] {
        arg_check(n >= 0 && (n == 0 || (P && lab && typ && part)), "seeding: null output array");
        DevMem mem;
        const Seeds s = run_seeding(g, grade, prm, pick_device(device), mem);
        arg_check(s.n == n, "seeding: the output holds " + std::to_string(n) + " nodes, the seeding made " +
                                std::to_string(s.n) + " (tn_seed_count)");
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

int tn_thin_nodes(const tn_grid_view* g, const tn_relax_params* prm, int64_t n, float* P, uint16_t* lab, uint8_t* typ,
                  uint16_t* part, uint16_t* part3, int64_t* n_kept, int32_t device) {
    return guarded([&] {
        arg_check(prm != nullptr && n_kept != nullptr, "thin: null argument");
        arg_check(n >= 0 && n < INT32_MAX && (n == 0 || (P && lab && typ && part)), "thin: bad node arrays");
        check_grid(g);
        *n_kept = n;
        if (!(prm->thin > 0.0f) || n == 0) {   // src/v2m_particles.cpp:1227-1229
            return;
        }
        *n_kept = thin_on_device(g, prm->thin, (int)n, P, lab, typ, part, part3, pick_device(device));
    });
}

int tn_protect_features(const tn_grid_view* g, const float* feat, int64_t nfeat, int64_t n, float* P, uint16_t* lab,
                        uint8_t* typ, uint16_t* part, uint16_t* part3, int64_t* n_kept, int32_t device) {
    return guarded([&] {
        arg_check(n_kept != nullptr && nfeat >= 0 && (nfeat == 0 || feat), "features: null argument");
        arg_check(n >= 0 && n < INT32_MAX && (n == 0 || (P && lab && typ && part)), "features: bad node arrays");
        check_grid(g);
        *n_kept = n;
        if (nfeat == 0 || n == 0) {   // src/v2m_particles.cpp:1005-1007
            return;
        }
        *n_kept = protect_on_device(g, feat, nfeat, (int)n, P, lab, typ, part, part3, pick_device(device));
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

int tn_relax_noalloc(const tn_grid_view* g, const tn_relax_params* prm, tn_nodes_view* nd, tn_relax_stats* st,
                     int32_t device) {
    return guarded([&] {
        check_grid(g);
        check_params(prm);
        check_nodes(nd);
        const int dev = pick_device(device);
        TN_HIP(hipSetDevice(dev));
        ScratchSlot& slot = scratch_slot(dev);
        std::lock_guard<std::mutex> hold(slot.lock);
        void* base = scratch_base(slot);
        // the array keeps the last call's contents: zero it (a fill kernel on the null
        // stream, ahead of every copy and kernel of this call)
        const int64_t words = (int64_t)(TN_NOALLOC_SCRATCH_BYTES / sizeof(int32_t));
        tnk::k_fill_i32<<<blocks_for(words, tnk::kBlock), tnk::kBlock, 0, nullptr>>>(static_cast<int32_t*>(base),
                                                                                      words, 0);
        TN_HIP(hipGetLastError());
        int64_t nkeys = 0;
        const tnd::Hash H = make_hash(g, prm, &nkeys);
        std::unique_ptr<tn_plan> p(plan_build(prm, nd, dev, H, nkeys, [&](DevMem& mem) { return upload_grid(g, mem); },
                                              base, TN_NOALLOC_SCRATCH_BYTES));
        if (p->n > 0) {   // plan_build left it reset: seeds in place, first rebuild done
            p->run_iters(p->prm.max_iters);
        }
        p->fill_stats(st);
        const int rc = tn_plan_download(p.get(), nd);
        if (rc != TN_OK) {
            throw TnError(rc, g_err);
        }
    });
}

int tn_relax_batch(const tn_grid_view* grids, const tn_relax_params* prms, tn_nodes_view* nodes,
                   tn_relax_stats* stats, int32_t count, int32_t device) {
    return guarded([&] {
        arg_check(count >= 0 && (count == 0 || (grids && prms && nodes)), "batch: null arrays");
        const auto t_start = std::chrono::steady_clock::now();
        struct Destroy {
            void operator()(tn_plan* p) const {
                tn_plan_destroy(p);
            }
        };
        std::vector<std::unique_ptr<tn_plan, Destroy>> plans;
        plans.reserve(count);
        for (int k = 0; k < count; ++k) {   // each plan: grid and seeds uploaded, first rebuild done
            tn_plan* p = nullptr;
            const int rc = tn_plan_create(&grids[k], &prms[k], &nodes[k], device, &p);
            if (rc != TN_OK) {
                throw TnError(rc, "volume " + std::to_string(k) + ": " + g_err);
            }
            plans.emplace_back(p);
        }
        const auto t_created = std::chrono::steady_clock::now();
        int64_t launches = 0, steps = 0;
        if (count > 0) {
            run_grouped(plans, launches, steps);
        }
        const auto t_looped = std::chrono::steady_clock::now();
        for (int k = 0; k < count; ++k) {
            plans[k]->fill_stats(stats ? &stats[k] : nullptr);
        }
        for (int k = 0; k < count; ++k) {
            const int rc = tn_plan_download(plans[k].get(), &nodes[k]);
            if (rc != TN_OK) {
                throw TnError(rc, "volume " + std::to_string(k) + ": " + g_err);
            }
        }
        if (std::getenv("TN_BATCH_VERBOSE")) {
            const auto ms = [](auto a, auto b) { return std::chrono::duration<double, std::milli>(b - a).count(); };
            std::fprintf(stderr, "[tn_relax_batch] %d volumes: plans %.1f ms, loop %.1f ms (%lld grouped iterations, "
                         "%lld volume-iterations), download %.1f ms\n", count, ms(t_start, t_created),
                         ms(t_created, t_looped), (long long)launches, (long long)steps,
                         ms(t_looped, std::chrono::steady_clock::now()));
        }
    });
}

}  // extern "C"

#include "tn_adjoint.hpp"

#include "tn_grid.hpp"
