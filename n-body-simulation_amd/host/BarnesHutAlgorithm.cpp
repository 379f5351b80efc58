#include "BarnesHutAlgorithm.hpp"

BarnesHutAlgorithm::BarnesHutAlgorithm(double dt, double tEnd, double visualizationStepWidth,
                                       std::string &outputDirectory)
    : nBodyAlgorithm(dt, tEnd, visualizationStepWidth, outputDirectory) {
    description = "Barnes-Hut Algorithm";
}

void BarnesHutAlgorithm::startSimulation(const SimulationData &simulationData) {
    runSimulation(simulationData, NB_ALGO_BARNES_HUT);
}
