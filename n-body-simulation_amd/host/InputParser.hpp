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
    // Throws std::invalid_argument on an unreadable file, a line with fewer
    // than 10 fields or a field with no leading number, and std::out_of_range
    // on a number std::stod cannot represent: the exceptions of the
    // reference's std::stod, thrown for the first bad line in file order.
    // The file is parsed in line-aligned pieces on the host's worker threads.
    static void parse_input(std::string &path, SimulationData &simulationData);

    // parse_input on the text of a file (header line included), cut into
    // `pieces` pieces parsed on up to `threads` threads; the result does not
    // depend on either.  `source` names the input in error messages.
    static void parse_text(const std::string &text, const std::string &source,
                           SimulationData &simulationData, unsigned threads, size_t pieces);

    // Splits at every ',' and keeps empty fields (InputParser.cpp:34-50).
    static std::vector<std::string> splitString(std::string string);
};

#endif
