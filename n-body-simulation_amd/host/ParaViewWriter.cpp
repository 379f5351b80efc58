#include "ParaViewWriter.hpp"

#include <charconv>
#include <cstdio>
#include <cstring>
#include <map>
#include <stdexcept>
#include <string_view>

#include "Parallel.hpp"

namespace paraview {
namespace {

// Appends text to a growing buffer without the per-call overhead of a stream.
class Out {
   public:
    explicit Out(std::string &s) : s_(s) {}
    void lit(std::string_view t) { s_.append(t.data(), t.size()); }
    void ch(char c) { s_.push_back(c); }
    // ostream << double with default flags: "%.6g"
    void g6(double v) {
        char buf[32];
        const auto r = std::to_chars(buf, buf + sizeof buf, v, std::chars_format::general, 6);
        s_.append(buf, (size_t)(r.ptr - buf));
    }
    template <class I>
    void integer(I v) {
        char buf[24];
        const auto r = std::to_chars(buf, buf + sizeof buf, v);
        s_.append(buf, (size_t)(r.ptr - buf));
    }

   private:
    std::string &s_;
};

// Orbit class codes of https://pdssbn.astro.umd.edu/data_other/objclass.shtml
// plus stars, dwarf planets, planets and moons; 0 for anything else
// (nBodyAlgorithm.cpp:298-342).
int classCode(const std::string &c) {
    static const std::map<std::string, int> code = {
        {"AMO", 1},  {"APO", 2},  {"ATE", 3},  {"IEO", 4},  {"MCA", 5},  {"IMB", 6},
        {"MBA", 7},  {"OMB", 8},  {"CEN", 9},  {"TJN", 10}, {"TNO", 11}, {"AST", 12},
        {"PAA", 13}, {"HYA", 14}, {"STA", 15}, {"DWA", 16}, {"PLA", 17}, {"SAT", 18}};
    auto it = code.find(c);
    return it != code.end() ? it->second : 0;
}

// The per-body sections of a .vtp file, in file order.
enum Section { kPos, kIds, kVel, kAcc, kMass, kNames, kClasses, kOffsets, kConn, kSections };

void formatChunk(const StoredStep &s, const SimulationData &d, size_t b, size_t e,
                 std::string *sec) {
    const size_t m = e - b;
    sec[kPos].reserve(m * 42);
    sec[kVel].reserve(m * 42);
    sec[kAcc].reserve(m * 14);
    sec[kMass].reserve(m * 14);
    sec[kIds].reserve(m * 9);
    sec[kOffsets].reserve(m * 9);
    sec[kConn].reserve(m * 9);
    sec[kClasses].reserve(m * 3);
    Out pos(sec[kPos]), ids(sec[kIds]), vel(sec[kVel]), acc(sec[kAcc]), mass(sec[kMass]),
        names(sec[kNames]), cls(sec[kClasses]), off(sec[kOffsets]), conn(sec[kConn]);
    for (size_t j = b; j < e; ++j) {
        pos.g6(s.px[j]), pos.ch(' '), pos.g6(s.py[j]), pos.ch(' '), pos.g6(s.pz[j]), pos.ch('\n');
        ids.integer((unsigned)j), ids.ch('\n');
        vel.g6(s.vx[j]), vel.ch(' '), vel.g6(s.vy[j]), vel.ch(' '), vel.g6(s.vz[j]), vel.ch('\n');
        acc.g6(s.acc[j]), acc.ch('\n');
        off.integer((unsigned)j + 1), off.ch(' ');
        conn.integer((unsigned)j), conn.ch(' ');
    }
    // masses, names and classes run over the input's bodies, which are the
    // same bodies in the same order
    const size_t ne = std::min(e, d.mass.size());
    for (size_t j = b; j < ne; ++j) mass.g6(d.mass[j]), mass.ch('\n');
    for (size_t j = b; j < std::min(e, d.names.size()); ++j) {
        for (char c : d.names[j]) names.integer((int)c), names.ch(' ');
        names.lit(" 0\n");
    }
    for (size_t j = b; j < std::min(e, d.body_classes.size()); ++j)
        cls.integer(classCode(d.body_classes[j])), cls.ch('\n');
}

size_t chunksOf(size_t n, size_t chunk) { return n == 0 ? 1 : (n + chunk - 1) / chunk; }

// Formats the per-body sections of one file: sec[k * chunks + c].
std::vector<std::string> formatSections(const StoredStep &s, const SimulationData &d,
                                        unsigned threads, size_t chunk) {
    if (chunk == 0) throw std::invalid_argument("paraview: chunk size must be positive");
    const size_t chunks = chunksOf(s.n, chunk);
    std::vector<std::string> per(chunks * kSections);
    host_parallel::parallelFor(chunks, threads, [&](size_t c) {
        std::string sec[kSections];
        formatChunk(s, d, c * chunk, std::min(s.n, (c + 1) * chunk), sec);
        for (int k = 0; k < kSections; ++k) per[(size_t)k * chunks + c] = std::move(sec[k]);
    });
    return per;
}

// The pieces of a .vtp file in order; the large ones are moved, not copied.
std::vector<std::string> vtpPieces(const StoredStep &s, const SimulationData &d, unsigned threads,
                                   size_t chunk) {
    std::vector<std::string> sec = formatSections(s, d, threads, chunk);
    const size_t chunks = chunksOf(s.n, chunk);
    std::vector<std::string> p;
    p.reserve(sec.size() + 32);
    auto section = [&](Section k) {
        for (size_t c = 0; c < chunks; ++c) p.push_back(std::move(sec[(size_t)k * chunks + c]));
    };
    std::string head;
    Out h(head);
    h.lit("<?xml version=\"1.0\"?>\n"
          "<VTKFile type=\"PolyData\" version=\"0.1\" byte_order=\"LittleEndian\" "
          "header_type=\"UInt64\">\n<PolyData>\n<Piece NumberOfPoints=\"");
    h.integer(s.n);
    h.lit("\" NumberOfVerts=\"");
    h.integer(s.n);
    h.lit("\">\n<Points>\n"
          "<DataArray type=\"Float64\" Name=\"position\" NumberOfComponents=\"3\" "
          "format=\"ascii\">\n");
    p.push_back(std::move(head));
    section(kPos);
    p.emplace_back("</DataArray>\n</Points>\n<PointData>\n"
                   "<DataArray type=\"Int32\" Name=\"body_id\" NumberOfComponents=\"1\" "
                   "format=\"ascii\">\n");
    section(kIds);
    p.emplace_back("</DataArray>\n"
                   "<DataArray type=\"Float64\" Name=\"velocity\" NumberOfComponents=\"3\" "
                   "format=\"ascii\">\n");
    section(kVel);
    p.emplace_back("</DataArray>\n"
                   "<DataArray type=\"Float64\" Name=\"acceleration\" NumberOfComponents=\"1\" "
                   "format=\"ascii\">\n");
    section(kAcc);
    p.emplace_back("</DataArray>\n"
                   "<DataArray type=\"Float64\" Name=\"mass\" NumberOfComponents=\"1\" "
                   "format=\"ascii\">\n");
    section(kMass);
    p.emplace_back("</DataArray>\n"
                   "<DataArray type=\"String\" Name=\"name\" NumberOfComponents=\"1\" "
                   "format=\"ascii\">\n");
    section(kNames);
    p.emplace_back("</DataArray>\n"
                   "<DataArray type=\"Int32\" Name=\"orbit_class\" NumberOfComponents=\"1\" "
                   "format=\"ascii\">\n");
    section(kClasses);
    p.emplace_back("</DataArray>\n</PointData>\n<Verts>\n"
                   "<DataArray type=\"Int64\" Name=\"offsets\">\n");
    section(kOffsets);
    p.emplace_back("\n</DataArray>\n<DataArray type=\"Int64\" Name=\"connectivity\">\n");
    section(kConn);
    std::string tail;
    Out t(tail);
    auto field = [&](const char *name, double v) {
        t.lit("<DataArray type=\"Float64\" Name=\"");
        t.lit(name);
        t.lit("\" NumberOfTuples=\"1\" format=\"ascii\">\n");
        t.g6(v);
        t.lit("\n</DataArray>\n");
    };
    t.lit("\n</DataArray>\n</Verts>\n</Piece>\n<FieldData>\n");
    field("kinetic energy", s.kinetic);
    field("potential energy", s.potential);
    field("total energy", s.total);
    field("virial equilibrium", s.virial);
    t.lit("</FieldData>\n</PolyData>\n</VTKFile>\n");
    p.push_back(std::move(tail));
    return p;
}

std::vector<std::string> lastStatePieces(const StoredStep &s, unsigned threads, size_t chunk) {
    if (chunk == 0) throw std::invalid_argument("paraview: chunk size must be positive");
    const size_t chunks = chunksOf(s.n, chunk);
    std::vector<std::string> p(chunks + 1);
    p[0] = "position_x, position_y, position_z \n";
    host_parallel::parallelFor(chunks, threads, [&](size_t c) {
        std::string &text = p[c + 1];
        const size_t b = c * chunk, e = std::min(s.n, (c + 1) * chunk);
        text.reserve((e - b) * 42);
        Out o(text);
        for (size_t j = b; j < e; ++j)
            o.g6(s.px[j]), o.ch(','), o.g6(s.py[j]), o.ch(','), o.g6(s.pz[j]), o.ch('\n');
    });
    return p;
}

std::string join(const std::vector<std::string> &p) {
    size_t total = 0;
    for (const auto &x : p) total += x.size();
    std::string r;
    r.reserve(total);
    for (const auto &x : p) r += x;
    return r;
}

void writeFile(const std::string &path, const std::vector<std::string> &p) {
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f) throw std::runtime_error("cannot open " + path + " for writing");
    bool ok = true;
    for (const auto &x : p)
        if (!x.empty() && std::fwrite(x.data(), 1, x.size(), f) != x.size()) ok = false;
    if (std::fclose(f) != 0) ok = false;
    if (!ok) throw std::runtime_error("write to " + path + " failed");
}

}  // namespace

std::string formatVtp(const StoredStep &s, const SimulationData &d, unsigned threads,
                      size_t chunk) {
    return join(vtpPieces(s, d, threads, chunk));
}

std::string formatLastState(const StoredStep &s, unsigned threads, size_t chunk) {
    return join(lastStatePieces(s, threads, chunk));
}

std::string formatPvd(size_t count) {
    std::string r;
    Out o(r);
    o.lit("<?xml version=\"1.0\"?>\n"
          "<VTKFile type=\"Collection\" version=\"0.1\" byte_order=\"LittleEndian\" "
          "compressor=\"vtkZLibDataCompressor\">\n<Collection>\n");
    for (size_t i = 0; i < count; ++i) {
        o.lit("<DataSet timestep=\"");
        o.integer(i);
        o.lit("\" group=\"\" part=\"0\" file=\"simulation_step");
        o.integer(i);
        o.lit(".vtp\"/>\n");
    }
    o.lit("</Collection>\n</VTKFile>\n");
    return r;
}

void writeSeries(const std::string &dir, const std::vector<StoredStep> &steps,
                 const SimulationData &d, unsigned threads, size_t chunk) {
    const std::string base = dir.empty() || dir.back() == '/' ? dir : dir + '/';
    if (!steps.empty()) writeFile(base + "lastState.csv", lastStatePieces(steps.back(), threads, chunk));
    // the reference opens the collection as "<dir>//simulation.pvd"
    writeFile(base + "/simulation.pvd", {formatPvd(steps.size())});
    auto vtp = [&](size_t i, unsigned t) {
        writeFile(base + "simulation_step" + std::to_string(i) + ".vtp",
                  vtpPieces(steps[i], d, t, chunk));
    };
    const bool oneChunk = steps.empty() || steps.front().n <= chunk;
    if (oneChunk) {
        // small files: one file per thread
        host_parallel::parallelFor(steps.size(), threads, [&](size_t i) { vtp(i, 1); });
    } else {
        for (size_t i = 0; i < steps.size(); ++i) vtp(i, threads);
    }
}

}  // namespace paraview
