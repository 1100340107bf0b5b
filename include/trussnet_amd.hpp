// trussnet_amd.hpp -- C++ wrapper of the C ABI (trussnet_amd.h).
//
// Errors become std::runtime_error, the exception v2mesh's relax_cl throws
// (src/v2m_gpu.h:16-17) and its pipeline catches to fall back to the CPU path
// (src/v2m_pipeline.cpp:497-526). Header-only; link libtrussnet_amd.so.
#ifndef TRUSSNET_AMD_HPP
#define TRUSSNET_AMD_HPP

#include <stdexcept>
#include <string>
#include <utility>

#include "trussnet_amd.h"

namespace tnamd {

// A failed call: the status code and the library's message.
class Error : public std::runtime_error {
public:
    Error(int code, const std::string& what) : std::runtime_error(what), code_(code) {}
    int code() const noexcept {
        return code_;
    }

private:
    int code_;
};

inline void check(int rc, const char* call) {
    if (rc != TN_OK) {
        throw Error(rc, std::string(call) + ": " + tn_last_error());
    }
}

inline tn_relax_params default_params() {
    tn_relax_params p;
    tn_relax_params_default(&p);
    return p;
}

// relax_cl's contract: relax the nodes in place on HIP device `device`, fill st.
inline void relax(const tn_grid_view& g, const tn_relax_params& prm, tn_nodes_view& nd, tn_relax_stats& st,
                  int device = -1) {
    check(tn_relax(&g, &prm, &nd, &st, device), "tn_relax");
}

// A resident relaxation plan (tn_plan_*), released on destruction.
class Plan {
public:
    Plan(const tn_grid_view& g, const tn_relax_params& prm, const tn_nodes_view& seeds, int device = -1) {
        check(tn_plan_create(&g, &prm, &seeds, device, &p_), "tn_plan_create");
    }
    Plan(const Plan&) = delete;
    Plan& operator=(const Plan&) = delete;
    Plan(Plan&& o) noexcept : p_(std::exchange(o.p_, nullptr)) {}
    ~Plan() {
        if (p_) {
            tn_plan_destroy(p_);
        }
    }
    tn_relax_stats run() {
        tn_relax_stats st;
        check(tn_plan_run(p_, &st), "tn_plan_run");
        return st;
    }
    void download(tn_nodes_view& nd) {
        check(tn_plan_download(p_, &nd), "tn_plan_download");
    }
    tn_plan* get() const noexcept {
        return p_;
    }

    // a plan over a device-built grid (tn_plan_create_from_grid)
    static Plan from_grid(const tn_grid* g, const tn_relax_params& prm, const tn_nodes_view& seeds) {
        Plan p;
        check(tn_plan_create_from_grid(g, &prm, &seeds, &p.p_), "tn_plan_create_from_grid");
        return p;
    }

private:
    Plan() = default;
    tn_plan* p_ = nullptr;
};

// A grid built on the device (tn_grid_build), released on destruction.
class DeviceGrid {
public:
    DeviceGrid(const tn_volume_view& vol, const tn_grid_params& prm, int device = -1) {
        check(tn_grid_build(&vol, &prm, device, &g_), "tn_grid_build");
    }
    DeviceGrid(const DeviceGrid&) = delete;
    DeviceGrid& operator=(const DeviceGrid&) = delete;
    DeviceGrid(DeviceGrid&& o) noexcept : g_(std::exchange(o.g_, nullptr)) {}
    ~DeviceGrid() {
        if (g_) {
            tn_grid_destroy(g_);
        }
    }
    tn_grid_info info() const {
        tn_grid_info in;
        check(tn_grid_get_info(g_, &in), "tn_grid_get_info");
        return in;
    }
    int64_t seed_count(const tn_relax_params& prm) const {
        int64_t n = 0;
        check(tn_seed_count_from_grid(g_, &prm, &n), "tn_seed_count_from_grid");
        return n;
    }
    // the Grid's arrays to the host (any pointer may be null; sizes in tn_grid_download)
    void download(int32_t* bl_cnt, uint16_t* bl_lab, int32_t* bl_slot, float* phi, float* h, uint8_t* grade,
                  float* gTW) const {
        check(tn_grid_download(g_, bl_cnt, bl_lab, bl_slot, phi, h, grade, gTW), "tn_grid_download");
    }
    const tn_grid* get() const noexcept {
        return g_;
    }

private:
    tn_grid* g_ = nullptr;
};

inline tn_grid_params default_grid_params() {
    tn_grid_params p;
    tn_grid_params_default(&p);
    return p;
}

}  // namespace tnamd

#endif  // TRUSSNET_AMD_HPP
