// BarnesHutAlgorithm.hpp -- Barnes-Hut simulation
// (src/simulationBackend/BarnesHutAlgorithm.hpp).  Each step rebuilds the
// octree (nb_bh_build) and traverses it (nb_bh_forces): with precision FP64
// the reference's per-body traversal, with FP32 the group traversal.
#ifndef NBODY_HOST_BARNESHUTALGORITHM_HPP
#define NBODY_HOST_BARNESHUTALGORITHM_HPP

#include "nBodyAlgorithm.hpp"

class BarnesHutAlgorithm : public nBodyAlgorithm {
   public:
    BarnesHutAlgorithm(double dt, double tEnd, double visualizationStepWidth,
                       std::string &outputDirectory);
    void startSimulation(const SimulationData &simulationData) override;
};

#endif
