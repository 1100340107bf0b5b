// v2m_hip.h -- v2mesh's node-placement entry points on an MI355X, through libtrussnet_amd.
//
// The source file a v2mesh maintainer adds to src/ (beside v2m_gpu.cpp) and links
// with -ltrussnet_amd. Each function keeps the contract of the reference function it
// stands for and throws std::runtime_error (tnamd::Error) on any failure, so
// run_pipeline's try/catch restores the seeds and takes the CPU path
// (src/v2m_pipeline.cpp:497-526):
//
//   relax_hip        tn::relax_cl  (src/v2m_gpu.h:18)         / tn::relax_cpu (src/v2m_particles.h:99)
//   relax_hip_batch  (none: one relax_cl call per volume)
//   seed_hip         tn::seed_cpu  (src/v2m_particles.h:98)
//   thin_hip         tn::thin_nodes (src/v2m_particles.h:101)
//   protect_hip      tn::protect_features (src/v2m_particles.h:105)
//   build_grid_hip   tn::build_grid_cpu (src/v2m_grid.h:97)
#ifndef V2MESH_HIP_H
#define V2MESH_HIP_H

#include <cstddef>
#include <vector>

#include "v2m_grid.h"
#include "v2m_particles.h"

namespace tn {

void relax_hip(const Grid& g, const RelaxParams& prm, Nodes& nd, RelaxStats& st, int device);
void relax_hip_batch(const std::vector<const Grid*>& g, const RelaxParams& prm, std::vector<Nodes*>& nd,
                     std::vector<RelaxStats>& st, int device);
void seed_hip(const Grid& g, const RelaxParams& prm, Nodes& nd, int device);
size_t thin_hip(const Grid& g, const RelaxParams& prm, Nodes& nd, int device);
size_t protect_hip(const Grid& g, Nodes& nd, int device);
void build_grid_hip(const LabelVolume& lv, const GridParams& p, Grid& g, int device);

}  // namespace tn

#endif  // V2MESH_HIP_H
