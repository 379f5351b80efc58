// InputParser.hpp -- CSV input (src/simulationData/InputParser.hpp).
// Columns: id,name,class,mass[kg],x,y,z[AU],vx,vy,vz[AU/d]; the first line is
// a header (InputParser.cpp:7-32, README.md:133-138).
#ifndef NBODY_HOST_INPUTPARSER_HPP
#define NBODY_HOST_INPUTPARSER_HPP

#include <string>
#include <vector>

#include "SimulationData.hpp"

class InputParser {
   public:
    // Throws std::invalid_argument on an unreadable file or a malformed
    // number (the reference's std::stod throws the same exception type).
    static void parse_input(std::string &path, SimulationData &simulationData);

    // Splits at every ',' and keeps empty fields (InputParser.cpp:34-50).
    static std::vector<std::string> splitString(std::string string);
};

#endif
