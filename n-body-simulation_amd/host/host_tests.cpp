// host_tests.cpp -- unit tests of the host-side units, carrying over the
// reference's tests/InputParserTest.cpp:4-32 and
// tests/TimeConverterTest.cpp:4-47 (googletest is not available; a minimal
// runner reports one line per check and exits non-zero on a failure), plus
// byte-for-byte checks of the parallel output writer against a serial
// std::ostream rendition of the same files.
//
//   host_tests                    run the checks
//   host_tests --bench-writer N   time both writers on one step of N bodies
//   host_tests --bench-parser F   time both parsers on the CSV file F
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <chrono>
#include <fstream>
#include <functional>
#include <iterator>
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

// The serial parser: the reference's getline / splitString / std::stod loop
// (InputParser.cpp:7-32) with a trailing '\r' dropped and empty lines skipped.
void serialParse(const std::string &text, const std::string &source, SimulationData &d) {
    std::istringstream in(text);
    std::string line;
    std::getline(in, line);
    size_t lineNo = 1;
    while (std::getline(in, line)) {
        ++lineNo;
        if (!line.empty() && line.back() == '\r') line.pop_back();
        if (line.empty()) continue;
        const std::vector<std::string> t = InputParser::splitString(line);
        if (t.size() < 10)
            throw std::invalid_argument("line " + std::to_string(lineNo) + " of " + source +
                                        ": expected 10 comma-separated fields");
        d.names.push_back(t[1]);
        d.body_classes.push_back(t[2]);
        d.mass.push_back(std::stod(t[3]));
        d.positions_x.push_back(std::stod(t[4]));
        d.positions_y.push_back(std::stod(t[5]));
        d.positions_z.push_back(std::stod(t[6]));
        d.velocities_x.push_back(std::stod(t[7]));
        d.velocities_y.push_back(std::stod(t[8]));
        d.velocities_z.push_back(std::stod(t[9]));
    }
}

bool sameBits(const std::vector<double> &a, const std::vector<double> &b) {
    return a.size() == b.size() && (a.empty() || std::memcmp(a.data(), b.data(), a.size() * 8) == 0);
}

bool sameData(const SimulationData &a, const SimulationData &b) {
    return a.names == b.names && a.body_classes == b.body_classes && sameBits(a.mass, b.mass) &&
           sameBits(a.positions_x, b.positions_x) && sameBits(a.positions_y, b.positions_y) &&
           sameBits(a.positions_z, b.positions_z) && sameBits(a.velocities_x, b.velocities_x) &&
           sameBits(a.velocities_y, b.velocities_y) && sameBits(a.velocities_z, b.velocities_z);
}

// "ok", or the exception's type and message
template <class F>
std::string outcome(F f) {
    try {
        f();
        return "ok";
    } catch (const std::invalid_argument &e) {
        return std::string("invalid_argument: ") + e.what();
    } catch (const std::out_of_range &e) {
        return std::string("out_of_range: ") + e.what();
    }
}

// A CSV whose numbers exercise std::stod: signs, exponents, hex floats,
// inf/nan, leading blanks, trailing junk after a number, 17-digit values;
// with CRLF and LF line ends, empty lines, extra fields and no final newline.
std::string makeCsv(size_t rows, uint64_t seed) {
    const char *odd[] = {" 1.5",    "+.5",  "-0",        "0x1p-3",  "inf",     "-INF",
                         "nan",     "1e5x", "7.25e+300", "  -2e-7", "1E3",     "0.000",
                         "12345678901234567890", "4.9406564584124654e-310",  // subnormal
                         "1e400",   "abc"};
    std::mt19937_64 rng(seed);
    std::uniform_real_distribution<double> u(-1.0, 1.0);
    std::string s = "id,name,class,mass,x,y,z,vx,vy,vz\n";
    char buf[40];
    for (size_t i = 0; i < rows; ++i) {
        s += std::to_string(i) + ",body " + std::to_string(i) + (i % 5 ? ",MBA" : ",");
        for (int k = 0; k < 7; ++k) {
            const uint64_t r = rng();
            if (r % 23 == 0) {
                const size_t which = (r >> 8) % 13;  // the 13 that std::stod accepts
                s += std::string(",") + odd[which];
            } else {
                std::snprintf(buf, sizeof buf, ",%.17g", u(rng) * std::pow(10.0, (int)(r % 40) - 20));
                s += buf;
            }
        }
        if (i % 11 == 3) s += ",extra,fields";
        s += i % 7 == 2 ? "\r\n" : "\n";
        if (i % 13 == 5) s += i % 2 ? "\n" : "\r\n";
    }
    s.pop_back();  // no final newline
    return s;
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

int benchParser(const std::string &path) {
    std::ifstream in(path, std::ios::binary);
    std::string text((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
    using clk = std::chrono::steady_clock;
    SimulationData a, b;
    auto t0 = clk::now();
    serialParse(text, path, a);
    auto t1 = clk::now();
    InputParser::parse_text(text, path, b, 8, 32);
    auto t2 = clk::now();
    std::printf("{\"rows\": %zu, \"bytes\": %zu, \"serial_s\": %.3f, \"parallel_8_threads_s\": %.3f, "
                "\"identical\": %s}\n",
                a.mass.size(), text.size(), std::chrono::duration<double>(t1 - t0).count(),
                std::chrono::duration<double>(t2 - t1).count(), sameData(a, b) ? "true" : "false");
    return sameData(a, b) ? 0 : 1;
}

int main(int argc, char **argv) {
    if (argc == 3 && std::string(argv[1]) == "--bench-parser") return benchParser(argv[2]);
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
    {
        // parallel parser == serial parser: values bit for bit, names, classes,
        // and the first error's type and message, for any piece count
        const std::string good = makeCsv(700, 3);
        SimulationData want;
        serialParse(good, "in.csv", want);
        bool ok = want.mass.size() == 700;
        for (unsigned t : {1u, 4u})
            for (size_t pieces : {(size_t)1, (size_t)2, (size_t)7, (size_t)64, (size_t)5000}) {
                SimulationData got;
                InputParser::parse_text(good, "in.csv", got, t, pieces);
                ok = ok && sameData(got, want);
            }
        check(ok, "InputParser.parallelMatchesSerial");
        // each error alone, placed on different lines; then two errors, where
        // the earlier line must win whatever the pieces
        const char *bad[] = {"1,a,b,1,2,3,4,5,6", "1,a,b,1,2,3,abc,5,6,7", "1,a,b,1e400,2,3,4,5,6,7",
                             "1,a,b,1,2,3,4,5,6,1e-320", "1,a,b,1,2,3,4,5,,7", ""};
        bool same = true;
        std::map<std::string, int> seen;
        for (const char *line : bad)
            for (size_t at : {(size_t)0, (size_t)350, (size_t)699}) {
                std::string text = good;
                size_t pos = 0;
                for (size_t l = 0; l <= at; ++l) pos = text.find('\n', pos) + 1;
                text.insert(pos, std::string(line) + "\n");
                std::string text2 = text;
                pos = 0;
                for (size_t l = 0; l <= at + 200 && pos != std::string::npos + 1; ++l)
                    pos = text2.find('\n', pos) + 1;
                if (pos != 0) text2.insert(pos, "9,x,y,oops,0,0,0,0,0,0\n");
                for (const std::string *tx : {&text, &text2}) {
                    SimulationData a;
                    const std::string w = outcome([&] { serialParse(*tx, "in.csv", a); });
                    ++seen[w.substr(0, w.find(' ', 22))];
                    for (size_t pieces : {(size_t)1, (size_t)3, (size_t)40, (size_t)999}) {
                        SimulationData b;
                        const std::string g = outcome(
                            [&] { InputParser::parse_text(*tx, "in.csv", b, 4, pieces); });
                        if (g != w || (w == "ok" && !sameData(a, b))) {
                            same = false;
                            std::printf("  mismatch at %zu pieces %zu: '%s' vs '%s'\n", at,
                                        pieces, g.c_str(), w.c_str());
                        }
                    }
                }
            }
        check(same && seen.size() >= 4, "InputParser.firstErrorMatchesSerial");
        for (auto &kv : seen) std::printf("  outcome %s x%d\n", kv.first.c_str(), kv.second);
        SimulationData e1, e2;
        InputParser::parse_text("", "in.csv", e1, 4, 8);
        InputParser::parse_text("id,name\n", "in.csv", e2, 4, 8);
        check(e1.mass.empty() && e2.mass.empty(), "InputParser.headerOnly");
    }
    std::printf("%d failure(s)\n", failures);
    return failures == 0 ? 0 : 1;
}
