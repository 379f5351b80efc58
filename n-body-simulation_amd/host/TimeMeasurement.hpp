// TimeMeasurement.hpp -- named timing sequences exported as times.json
// (src/utility/TimeMeasurement.{hpp,cpp}; same keys and layout).
#ifndef NBODY_HOST_TIMEMEASUREMENT_HPP
#define NBODY_HOST_TIMEMEASUREMENT_HPP

#include <map>
#include <string>
#include <vector>

#include "Configuration.hpp"

class TimeMeasurement {
   public:
    std::map<std::string, std::vector<double>> times;
    std::string algorithmType;
    std::string device;
    d_type::int_t bodyCount = 0;

    void addTimingSequence(const std::string &name);
    void addTimeToSequence(const std::string &sequenceName, double time);
    void exportJSON(const std::string &path);
    void setProperties(std::string &algorithm, d_type::int_t &bodyCountArg, std::string &deviceArg);
};

#endif
