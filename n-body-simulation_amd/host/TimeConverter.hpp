// TimeConverter.hpp -- "<double><h|d|m|y>" to days (src/utility/TimeConverter.cpp:4-48).
#ifndef NBODY_HOST_TIMECONVERTER_HPP
#define NBODY_HOST_TIMECONVERTER_HPP

#include <string>

class TimeConverter {
   public:
    // h: /24, d: 1, m: 30.4167, y: 365.25.  Throws std::invalid_argument when
    // the suffix is missing or unknown or the prefix is not entirely a number.
    static double convertToEarthDays(std::string &time);
};

#endif
