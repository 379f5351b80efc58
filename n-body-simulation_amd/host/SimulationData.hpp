// SimulationData.hpp -- the input of a simulation, one entry per body
// (same fields as the reference's src/simulationData/SimulationData.hpp:13-27).
#ifndef NBODY_HOST_SIMULATIONDATA_HPP
#define NBODY_HOST_SIMULATIONDATA_HPP

#include <string>
#include <vector>

struct SimulationData {
    std::vector<std::string> names;
    std::vector<std::string> body_classes;
    std::vector<double> mass;

    std::vector<double> positions_x;
    std::vector<double> positions_y;
    std::vector<double> positions_z;

    std::vector<double> velocities_x;
    std::vector<double> velocities_y;
    std::vector<double> velocities_z;
};

#endif
