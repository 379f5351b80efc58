// NaiveAlgorithm.hpp -- direct-sum simulation (src/simulationBackend/NaiveAlgorithm.hpp).
// Forces: nb_forces_direct (FP64 in the reference's summation order by
// default).  blockSize and optimization_stage are accepted for
// compatibility: every stage of the reference performs the same arithmetic.
#ifndef NBODY_HOST_NAIVEALGORITHM_HPP
#define NBODY_HOST_NAIVEALGORITHM_HPP

#include "nBodyAlgorithm.hpp"

class NaiveAlgorithm : public nBodyAlgorithm {
   public:
    NaiveAlgorithm(double dt, double tEnd, double visualizationStepWidth,
                   std::string &outputDirectory);
    void startSimulation(const SimulationData &simulationData) override;
};

#endif
