// host_tests.cpp -- unit tests of the host-side parsing units, carrying over
// the reference's tests/InputParserTest.cpp:4-32 and
// tests/TimeConverterTest.cpp:4-47 (googletest is not available; a minimal
// runner reports one line per check and exits non-zero on a failure).
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <stdexcept>
#include <string>
#include <vector>

#include "InputParser.hpp"
#include "TimeConverter.hpp"

namespace {

int failures = 0;

void check(bool ok, const std::string &what) {
    std::printf("%s %s\n", ok ? "PASS" : "FAIL", what.c_str());
    if (!ok) ++failures;
}

// googletest's EXPECT_FLOAT_EQ: the two values rounded to float lie within
// 4 units in the last place
bool float_eq(double a, double b) {
    float fa = (float)a, fb = (float)b;
    int32_t ia, ib;
    std::memcpy(&ia, &fa, 4);
    std::memcpy(&ib, &fb, 4);
    // map sign-magnitude to a monotone integer line
    const int64_t la = ia < 0 ? -(int64_t)(ia & 0x7fffffff) : ia;
    const int64_t lb = ib < 0 ? -(int64_t)(ib & 0x7fffffff) : ib;
    return std::llabs(la - lb) <= 4;
}

void expect_throw(const std::function<void()> &f, const std::string &what) {
    try {
        f();
    } catch (const std::invalid_argument &) {
        check(true, what);
        return;
    } catch (...) {
    }
    check(false, what);
}

}  // namespace

int main() {
    {
        auto r = InputParser::splitString("5,Name,Class,5000,0.5,-0.3,0.4,0.06,0.1,0.2");
        const std::vector<std::string> e = {"5", "Name", "Class", "5000", "0.5",
                                            "-0.3", "0.4", "0.06", "0.1", "0.2"};
        check(r == e && r.size() == 10, "SplitStringTest.BasicCase");
    }
    {
        auto r = InputParser::splitString("5,,Class,5000,0.5,-0.3,0.4,0.06,0.1,0.2");
        check(r.size() == 10 && r[1].empty() && r[2] == "Class" && r[9] == "0.2",
              "SplitStringTest.EmptyEntry");
    }
    auto conv = [](std::string s) { return TimeConverter::convertToEarthDays(s); };
    check(float_eq(conv("1h"), 1.0 / 24), "TimeConversionTests.convertHour1");
    check(float_eq(conv("10h"), 10.0 / 24), "TimeConversionTests.convertHour2");
    check(float_eq(conv("1m"), 30.4167), "TimeConversionTests.convertMonth");
    check(float_eq(conv("1y"), 365.25), "TimeConversionTests.convertYear");
    check(float_eq(conv("0.5y"), 182.625), "TimeConversionTests.convertYear3");
    check(conv("2d") == 2.0, "TimeConversionTests.convertDay");
    expect_throw([&] { conv("0.5s"); }, "TimeConversionTests.invalidDescriptor");
    expect_throw([&] { conv("0.5"); }, "TimeConversionTests.noDescriptor");
    expect_throw([&] { conv("1y 5m"); }, "TimeConversionTests.invalidFormat");
    expect_throw([&] { conv(""); }, "TimeConversionTests.empty");
    std::printf("%d failure(s)\n", failures);
    return failures == 0 ? 0 : 1;
}
