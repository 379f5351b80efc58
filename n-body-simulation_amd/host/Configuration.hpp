// Configuration.hpp -- global run configuration of the drop-in application.
//
// Same names, defaults and sizing rule as the reference's
// src/utility/Configuration.hpp:8-124 / Configuration.cpp:5-33.  The
// work-item counts only steer the reference's SYCL launches; they are kept,
// printed and written to times.json so that runs stay comparable, but the
// gfx950 kernels choose their own launch shapes.  `precision` is new: FP64
// (the reference's arithmetic, default) or FP32 pair forces with FP64
// accumulation.
#ifndef NBODY_HOST_CONFIGURATION_HPP
#define NBODY_HOST_CONFIGURATION_HPP

#include <cstdint>

namespace d_type {
typedef unsigned int int_t;  // Configuration.hpp:9
}

namespace configuration {
extern d_type::int_t numberOfBodies;
extern double epsilon2;
extern bool compute_energy;
extern bool use_GPUs;
extern int precision;  // NB_PREC_FP64 or NB_PREC_FP32C

namespace naive_algorithm {
extern int blockSize;
extern int optimization_stage;
}  // namespace naive_algorithm

namespace barnes_hut_algorithm {
extern double theta;
extern int workGroupSize;
extern d_type::int_t stackSize;
extern d_type::int_t storageSizeParameter;
extern int AABBWorkItemCount;
extern int octreeWorkItemCount;
extern int octreeTopWorkItemCount;
extern int centerOfMassWorkItemCount;
extern int maxBuildLevel;
extern bool sortBodies;
}  // namespace barnes_hut_algorithm

// Configuration.cpp:24-33
void initializeConfigValues(d_type::int_t bodyCount, int storageSizeParam, int stackSizeParam);

void setBlockSize(int blockSize);
void setTheta(double theta);
void setAABBWorkItemCount(int workItemCount);
void setOctreeWorkItemCount(int workItemCount);
void setOctreeTopWorkItemCount(int workItemCount);
void setCenterOfMassWorkItemCount(int workItemCount);
void setMaxBuildLevel(int maxLevel);
void setEnergyComputation(bool computeEnergy);
void setSortBodies(bool sort_bodies);
void setDeviceGPU(bool useGPU);
void setWorkGroupSizeBarnesHut(int workGroupSize);
void setOptimizationStage(int stage);
void setPrecision(int precision);
}  // namespace configuration

#endif
