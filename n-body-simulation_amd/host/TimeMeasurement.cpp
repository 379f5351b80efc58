#include "TimeMeasurement.hpp"

#include <fstream>

#include "nbody_capi.h"

void TimeMeasurement::addTimingSequence(const std::string &name) { times[name] = {}; }

void TimeMeasurement::addTimeToSequence(const std::string &sequenceName, double time) {
    times[sequenceName].push_back(time);
}

// Layout of TimeMeasurement.cpp:16-62 (keys in map order, sequences as
// arrays, default stream precision); "precision" is added after "device".
void TimeMeasurement::exportJSON(const std::string &path) {
    std::ofstream jsonFile(path);
    jsonFile << "{ \n";
    jsonFile << "  \"algorithm\": \"" << algorithmType << "\",\n";
    jsonFile << "  \"device\": \"" << device << "\",\n";
    jsonFile << "  \"precision\": \""
             << (configuration::precision == NB_PREC_FP64 ? "fp64" : "fp32+fp64 accumulation")
             << "\",\n";
    if (algorithmType == "Naive Algorithm") {
        jsonFile << "  \"block size\": " << configuration::naive_algorithm::blockSize << ",\n";
        jsonFile << "  \"optimization stage\": " << configuration::naive_algorithm::optimization_stage
                 << ",\n";
    } else {
        namespace bh = configuration::barnes_hut_algorithm;
        jsonFile << "  \"theta\": " << bh::theta << ",\n";
        jsonFile << "  \"work-group size acceleration\": " << bh::workGroupSize << ",\n";
        jsonFile << "  \"work-items AABB\": " << bh::AABBWorkItemCount << ",\n";
        jsonFile << "  \"work-items octree\": " << bh::octreeWorkItemCount << ",\n";
        jsonFile << "  \"work-items center of mass\": " << bh::centerOfMassWorkItemCount << ",\n";
        jsonFile << "  \"work-items top octree\": " << bh::octreeTopWorkItemCount << ",\n";
        jsonFile << "  \"max build-level top octree\": " << bh::maxBuildLevel << ",\n";
        jsonFile << "  \"bodies sorted\": " << bh::sortBodies << ",\n";
    }
    jsonFile << "  \"body count\": " << bodyCount;
    for (auto &[sequenceName, timeSequence] : times) {
        if (timeSequence.empty()) continue;
        jsonFile << ",\n";
        jsonFile << "  \"" << sequenceName << "\": [";
        for (size_t t = 0; t + 1 < timeSequence.size(); ++t) jsonFile << timeSequence[t] << ", ";
        jsonFile << timeSequence.back() << "]";
    }
    jsonFile << "\n}";
}

void TimeMeasurement::setProperties(std::string &algorithm, d_type::int_t &bodyCountArg,
                                    std::string &deviceArg) {
    algorithmType = algorithm;
    bodyCount = bodyCountArg;
    device = deviceArg;
}
