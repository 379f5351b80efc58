// host_tests.cpp -- unit tests of the host-side units, carrying over the
// reference's tests/InputParserTest.cpp:4-32 and
// tests/TimeConverterTest.cpp:4-47 (googletest is not available; a minimal
// runner reports one line per check and exits non-zero on a failure), plus
// byte-for-byte checks of the parallel output writer against a serial
// std::ostream rendition of the same files.
//
//   host_tests                    run the checks
//   host_tests --bench-writer N   time both writers on one step of N bodies
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <chrono>
#include <functional>
#include <limits>
#include <map>
#include <random>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "InputParser.hpp"
#include "ParaViewWriter.hpp"
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


// The .vtp text as the reference's stream writers produce it
// (nBodyAlgorithm.cpp:158-247, 286-395): every number through operator<<.
std::string serialVtp(const paraview::StoredStep &s, const SimulationData &d) {
    static const std::map<std::string, int> code = {
        {"AMO", 1},  {"APO", 2},  {"ATE", 3},  {"IEO", 4},  {"MCA", 5},  {"IMB", 6},
        {"MBA", 7},  {"OMB", 8},  {"CEN", 9},  {"TJN", 10}, {"TNO", 11}, {"AST", 12},
        {"PAA", 13}, {"HYA", 14}, {"STA", 15}, {"DWA", 16}, {"PLA", 17}, {"SAT", 18}};
    std::ostringstream v;
    const unsigned n = (unsigned)s.n;
    v << "<?xml version=\"1.0\"?>" << '\n'
      << R"(<VTKFile type="PolyData" version="0.1" byte_order="LittleEndian" header_type="UInt64">)"
      << '\n' << "<PolyData>" << '\n' << "<Piece NumberOfPoints=\"" << s.n
      << "\" NumberOfVerts=\"" << s.n << "\">" << '\n' << "<Points>" << '\n'
      << R"(<DataArray type="Float64" Name="position" NumberOfComponents="3" format="ascii">)"
      << '\n';
    for (unsigned j = 0; j < n; ++j) v << s.px[j] << " " << s.py[j] << " " << s.pz[j] << '\n';
    v << "</DataArray>" << '\n' << "</Points>" << '\n' << "<PointData>" << '\n'
      << R"(<DataArray type="Int32" Name="body_id" NumberOfComponents="1" format="ascii">)" << '\n';
    for (unsigned j = 0; j < n; ++j) v << j << '\n';
    v << "</DataArray>" << '\n'
      << R"(<DataArray type="Float64" Name="velocity" NumberOfComponents="3" format="ascii">)"
      << '\n';
    for (unsigned j = 0; j < n; ++j) v << s.vx[j] << " " << s.vy[j] << " " << s.vz[j] << '\n';
    v << "</DataArray>" << '\n'
      << R"(<DataArray type="Float64" Name="acceleration" NumberOfComponents="1" format="ascii">)"
      << '\n';
    for (unsigned j = 0; j < n; ++j) v << s.acc[j] << '\n';
    v << "</DataArray>" << '\n'
      << R"(<DataArray type="Float64" Name="mass" NumberOfComponents="1" format="ascii">)" << '\n';
    for (double m : d.mass) v << m << '\n';
    v << "</DataArray>" << '\n'
      << R"(<DataArray type="String" Name="name" NumberOfComponents="1" format="ascii">)" << '\n';
    for (const std::string &name : d.names) {
        for (char c : name) v << (int)c << ' ';
        v << " 0" << '\n';
    }
    v << "</DataArray>" << '\n'
      << R"(<DataArray type="Int32" Name="orbit_class" NumberOfComponents="1" format="ascii">)"
      << '\n';
    for (const std::string &c : d.body_classes) {
        auto it = code.find(c);
        v << (it != code.end() ? it->second : 0) << '\n';
    }
    v << "</DataArray>" << '\n' << "</PointData>" << '\n' << "<Verts>" << '\n'
      << R"(<DataArray type="Int64" Name="offsets">)" << '\n';
    for (unsigned j = 1; j <= n; ++j) v << std::to_string(j) << ' ';
    v << '\n' << "</DataArray>" << '\n' << R"(<DataArray type="Int64" Name="connectivity">)" << '\n';
    for (unsigned j = 0; j < n; ++j) v << std::to_string(j) << ' ';
    v << '\n' << "</DataArray>" << '\n' << "</Verts>" << '\n' << "</Piece>" << '\n'
      << "<FieldData>" << '\n';
    const char *fields[4] = {"kinetic energy", "potential energy", "total energy",
                             "virial equilibrium"};
    const double values[4] = {s.kinetic, s.potential, s.total, s.virial};
    for (int k = 0; k < 4; ++k)
        v << R"(<DataArray type="Float64" Name=")" << fields[k]
          << R"(" NumberOfTuples="1" format="ascii">)" << '\n' << values[k] << '\n'
          << "</DataArray>" << '\n';
    v << "</FieldData>" << '\n' << "</PolyData>" << '\n' << "</VTKFile>" << '\n';
    return v.str();
}

std::string serialLastState(const paraview::StoredStep &s) {
    std::ostringstream c;
    c << "position_x, position_y, position_z \n";
    for (size_t j = 0; j < s.n; ++j) c << s.px[j] << "," << s.py[j] << "," << s.pz[j] << '\n';
    return c.str();
}

// n bodies of values chosen to exercise "%.6g": signed zeros, subnormals,
// the extremes, infinities and NaNs, the fixed/exponent switch at 1e-5 and
// 1e6, halfway cases of the sixth digit, and random values over 600 decades.
struct WriterCase {
    std::vector<double> v[7];
    SimulationData d;
    paraview::StoredStep step;
};

void makeCase(WriterCase &w, size_t n, uint64_t seed) {
    const double special[] = {0.0, -0.0, 5e-324, -2.2250738585072014e-308, 1e-310,
                              std::numeric_limits<double>::max(), -1e308,
                              std::numeric_limits<double>::infinity(),
                              -std::numeric_limits<double>::infinity(),
                              std::numeric_limits<double>::quiet_NaN(),
                              -std::numeric_limits<double>::quiet_NaN(), 0.1, 1e-5, 9.99999e-6,
                              0.0001, 999999.5, 999999.4, 1e6, 123456.5, 9.9999995, 2.5e-7,
                              -1.234565, 1.0 / 3.0, 6.02214076e23, 4294967296.0, 1.5};
    const size_t ns = sizeof special / sizeof special[0];
    std::mt19937_64 rng(seed);
    std::uniform_real_distribution<double> mant(-10.0, 10.0);
    std::uniform_int_distribution<int> ex(-300, 300);
    for (auto &a : w.v) a.resize(n);
    for (size_t j = 0; j < n; ++j)
        for (int k = 0; k < 7; ++k)
            w.v[k][j] = (j * 7 + k) % 5 == 0 ? special[(j * 3 + k) % ns]
                                             : mant(rng) * std::pow(10.0, ex(rng));
    const char *classes[] = {"STA", "PLA", "MBA", "XYZ", "", "SAT", "TNO"};
    for (size_t j = 0; j < n; ++j) {
        w.d.mass.push_back(j % 3 == 0 ? special[j % ns] : std::fabs(mant(rng)) * 1e20);
        std::string name = j % 4 == 0 ? "" : "b" + std::to_string(j);
        if (j % 9 == 1) name += "\xc3\xa9";  // bytes above 127 print as negative codes
        w.d.names.push_back(name);
        w.d.body_classes.push_back(classes[j % 7]);
    }
    w.step.n = n;
    w.step.px = w.v[0].data(), w.step.py = w.v[1].data(), w.step.pz = w.v[2].data();
    w.step.vx = w.v[3].data(), w.step.vy = w.v[4].data(), w.step.vz = w.v[5].data();
    w.step.acc = w.v[6].data();
    w.step.kinetic = 1.2345678e-9, w.step.potential = -2.0 / 3.0, w.step.total = -0.0;
    w.step.virial = 1e300;
}

int benchWriter(size_t n) {
    WriterCase w;
    makeCase(w, n, 1);
    using clk = std::chrono::steady_clock;
    auto t0 = clk::now();
    const std::string a = serialVtp(w.step, w.d);
    auto t1 = clk::now();
    const unsigned threads = 8;
    const std::string b = paraview::formatVtp(w.step, w.d, threads);
    auto t2 = clk::now();
    const std::string c = paraview::formatVtp(w.step, w.d, 1);
    auto t3 = clk::now();
    const double ts = std::chrono::duration<double>(t1 - t0).count();
    const double tp = std::chrono::duration<double>(t2 - t1).count();
    const double t1s = std::chrono::duration<double>(t3 - t2).count();
    std::printf("{\"bodies\": %zu, \"bytes\": %zu, \"ostream_s\": %.4f, "
                "\"to_chars_1_thread_s\": %.4f, \"to_chars_%u_threads_s\": %.4f, "
                "\"identical\": %s}\n",
                n, a.size(), ts, t1s, threads, tp, (a == b && a == c) ? "true" : "false");
    return a == b && a == c ? 0 : 1;
}

}  // namespace

int main(int argc, char **argv) {
    if (argc == 3 && std::string(argv[1]) == "--bench-writer")
        return benchWriter((size_t)std::strtoull(argv[2], nullptr, 10));
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
    {
        // parallel writer == serial stream writer, for any threads and chunks
        for (size_t n : {0, 1, 5, 97, 1000}) {
            WriterCase w;
            makeCase(w, n, 42 + n);
            const std::string want = serialVtp(w.step, w.d);
            const std::string last = serialLastState(w.step);
            bool ok = true;
            for (unsigned t : {1u, 3u, 8u})
                for (size_t chunk : {(size_t)1, (size_t)7, (size_t)64, paraview::kChunkBodies}) {
                    ok = ok && paraview::formatVtp(w.step, w.d, t, chunk) == want;
                    ok = ok && paraview::formatLastState(w.step, t, chunk) == last;
                }
            check(ok, "ParaViewWriter.matchesStreamOutput n=" + std::to_string(n));
        }
        // simulation.pvd as nBodyAlgorithm.cpp:151-166, 248 streams it
        std::ostringstream pvd;
        pvd << "<?xml version=\"1.0\"?>" << '\n'
            << R"(<VTKFile type="Collection" version="0.1" byte_order="LittleEndian" compressor="vtkZLibDataCompressor">)"
            << '\n' << "<Collection>" << '\n';
        for (unsigned i = 0; i < 3; ++i)
            pvd << "<DataSet timestep=\"" << i << "\" " << R"(group="" part="0" file=")"
                << "simulation_step" + std::to_string(i) + ".vtp" << "\"/>" << '\n';
        pvd << "</Collection>" << '\n' << "</VTKFile>" << '\n';
        check(paraview::formatPvd(3) == pvd.str(),
              "ParaViewWriter.pvd");
    }
    std::printf("%d failure(s)\n", failures);
    return failures == 0 ? 0 : 1;
}
