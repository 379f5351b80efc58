#include "TimeConverter.hpp"

#include <stdexcept>

double TimeConverter::convertToEarthDays(std::string &time) {
    const double hourInEarthDays = 1.0 / 24;
    const double monthInEarthDays = 30.4167;
    const double yearInEarthDays = 365.25;
    const std::string message = "Time values have to be of the format <double><h|d|m|y>";
    if (time.empty()) throw std::invalid_argument(message);
    const char descriptor = time[time.length() - 1];
    const std::string prefix = time.substr(0, time.length() - 1);
    size_t used = 0;
    double value;
    try {
        value = std::stod(prefix, &used);
    } catch (const std::exception &) {
        throw std::invalid_argument(message);
    }
    if (used != prefix.length()) throw std::invalid_argument(message);
    switch (descriptor) {
        case 'h': return value * hourInEarthDays;
        case 'd': return value;
        case 'm': return value * monthInEarthDays;
        case 'y': return value * yearInEarthDays;
        default: throw std::invalid_argument(message);
    }
}
