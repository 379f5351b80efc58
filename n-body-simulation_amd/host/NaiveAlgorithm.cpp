#include "NaiveAlgorithm.hpp"

NaiveAlgorithm::NaiveAlgorithm(double dt, double tEnd, double visualizationStepWidth,
                               std::string &outputDirectory)
    : nBodyAlgorithm(dt, tEnd, visualizationStepWidth, outputDirectory) {
    description = "Naive Algorithm";
}

void NaiveAlgorithm::startSimulation(const SimulationData &simulationData) {
    runSimulation(simulationData, NB_ALGO_NAIVE);
}
