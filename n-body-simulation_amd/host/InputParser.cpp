#include "InputParser.hpp"

#include <fstream>
#include <stdexcept>

void InputParser::parse_input(std::string &path, SimulationData &simulationData) {
    std::ifstream in(path);
    if (!in) throw std::invalid_argument("cannot open input file " + path);
    std::string line;
    std::getline(in, line);  // header
    size_t lineNo = 1;
    while (std::getline(in, line)) {
        ++lineNo;
        if (!line.empty() && line.back() == '\r') line.pop_back();
        if (line.empty()) continue;
        const std::vector<std::string> tokens = splitString(line);
        if (tokens.size() < 10)
            throw std::invalid_argument("line " + std::to_string(lineNo) + " of " + path +
                                        ": expected 10 comma-separated fields");
        simulationData.names.push_back(tokens[1]);
        simulationData.body_classes.push_back(tokens[2]);
        // std::stod, as the reference: leading numeric prefix, throws on none
        simulationData.mass.push_back(std::stod(tokens[3]));
        simulationData.positions_x.push_back(std::stod(tokens[4]));
        simulationData.positions_y.push_back(std::stod(tokens[5]));
        simulationData.positions_z.push_back(std::stod(tokens[6]));
        simulationData.velocities_x.push_back(std::stod(tokens[7]));
        simulationData.velocities_y.push_back(std::stod(tokens[8]));
        simulationData.velocities_z.push_back(std::stod(tokens[9]));
    }
}

std::vector<std::string> InputParser::splitString(std::string string) {
    std::vector<std::string> result;
    size_t start = 0;
    for (size_t i = 0; i < string.size(); ++i) {
        if (string[i] == ',') {
            result.push_back(string.substr(start, i - start));
            start = i + 1;
        }
    }
    result.push_back(string.substr(start));
    return result;
}
