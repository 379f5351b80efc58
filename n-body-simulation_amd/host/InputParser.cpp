#include "InputParser.hpp"

#include <cstdio>
#include <exception>
#include <stdexcept>

#include "Parallel.hpp"

namespace {

// The rows of one line-aligned piece of the input, or the first error in it.
struct Piece {
    SimulationData rows;
    size_t lines = 0;               // lines that start in the piece, empty ones included
    size_t errorLine = 0;           // 1-based line within the piece of the first error
    std::exception_ptr error;
};

// std::stod on the field [b, e): the same conversion, the same exceptions
// (std::invalid_argument without a leading number, std::out_of_range when the
// value over- or underflows), as the reference's std::stod(tokens[k]).
double stodField(const char *b, const char *e, std::string &scratch) {
    scratch.assign(b, e);
    return std::stod(scratch);
}

// Parses lines [b, e) of `text` the way the reference's loop parses each
// line (InputParser.cpp:15-31), with two additions: a trailing '\r' is
// dropped and empty lines are skipped.
void parsePiece(const char *b, const char *e, const std::string &source, Piece &out) {
    std::string scratch;
    const char *f[11];  // starts of fields 0..9 and the end of field 9
    for (const char *p = b; p < e;) {
        const char *nl = p;
        while (nl < e && *nl != '\n') ++nl;
        const char *lineEnd = nl;
        const char *next = nl < e ? nl + 1 : e;
        ++out.lines;
        if (lineEnd > p && lineEnd[-1] == '\r') --lineEnd;
        if (lineEnd == p) {
            p = next;
            continue;
        }
        int nf = 1;
        f[0] = p;
        for (const char *c = p; c < lineEnd && nf <= 10; ++c)
            if (*c == ',') f[nf++] = c + 1;
        try {
            if (nf < 10)
                throw std::invalid_argument("line " + std::to_string(out.lines) + " of " + source +
                                            ": expected 10 comma-separated fields");
            // field k runs from f[k] to the comma before f[k + 1] (or the line end)
            auto end = [&](int k) { return k + 1 < nf ? f[k + 1] - 1 : lineEnd; };
            SimulationData &r = out.rows;
            r.names.emplace_back(f[1], end(1));
            r.body_classes.emplace_back(f[2], end(2));
            r.mass.push_back(stodField(f[3], end(3), scratch));
            r.positions_x.push_back(stodField(f[4], end(4), scratch));
            r.positions_y.push_back(stodField(f[5], end(5), scratch));
            r.positions_z.push_back(stodField(f[6], end(6), scratch));
            r.velocities_x.push_back(stodField(f[7], end(7), scratch));
            r.velocities_y.push_back(stodField(f[8], end(8), scratch));
            r.velocities_z.push_back(stodField(f[9], end(9), scratch));
        } catch (...) {
            out.errorLine = out.lines;
            out.error = std::current_exception();
            return;
        }
        p = next;
    }
}

template <class T>
void append(std::vector<T> &dst, std::vector<T> &src) {
    dst.insert(dst.end(), std::make_move_iterator(src.begin()), std::make_move_iterator(src.end()));
}

}  // namespace

void InputParser::parse_input(std::string &path, SimulationData &simulationData) {
    std::FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) throw std::invalid_argument("cannot open input file " + path);
    std::string text;
    char buf[1 << 16];
    size_t got;
    while ((got = std::fread(buf, 1, sizeof buf, f)) > 0) text.append(buf, got);
    const bool bad = std::ferror(f) != 0;
    std::fclose(f);
    if (bad) throw std::invalid_argument("cannot read input file " + path);
    // pieces of at least 64 KiB, four per thread
    const unsigned threads = host_parallel::defaultThreads();
    const size_t pieces = std::min<size_t>(4 * (size_t)threads, text.size() / 65536 + 1);
    parse_text(text, path, simulationData, threads, pieces);
}

void InputParser::parse_text(const std::string &text, const std::string &source,
                             SimulationData &simulationData, unsigned threads, size_t pieces) {
    // the first line is the header
    const size_t nl = text.find('\n');
    const size_t start = nl == std::string::npos ? text.size() : nl + 1;
    const size_t len = text.size() - start;
    pieces = std::max<size_t>(1, pieces);
    // piece k starts at the first line start at or after start + k * len / pieces
    std::vector<size_t> cut(pieces + 1);
    cut[0] = start;
    cut[pieces] = text.size();
    for (size_t k = 1; k < pieces; ++k) {
        size_t p = std::max(cut[k - 1], start + k * (len / pieces));
        if (p > start && p < text.size() && text[p - 1] != '\n') {
            const size_t q = text.find('\n', p);
            p = q == std::string::npos ? text.size() : q + 1;
        }
        cut[k] = p;
    }
    std::vector<Piece> out(pieces);
    host_parallel::parallelFor(pieces, threads, [&](size_t k) {
        parsePiece(text.data() + cut[k], text.data() + cut[k + 1], source, out[k]);
    });
    // the first error in file order, with the file's line number
    size_t lineBase = 1;  // the header
    for (Piece &p : out) {
        if (p.error) {
            try {
                std::rethrow_exception(p.error);
            } catch (const std::invalid_argument &e) {
                const std::string msg = e.what();
                if (msg.rfind("line ", 0) == 0)  // renumber from piece to file lines
                    throw std::invalid_argument("line " + std::to_string(lineBase + p.errorLine) +
                                                msg.substr(msg.find(' ', 5)));
                throw;
            }
        }
        lineBase += p.lines;
    }
    SimulationData &d = simulationData;
    for (Piece &p : out) {
        append(d.names, p.rows.names);
        append(d.body_classes, p.rows.body_classes);
        append(d.mass, p.rows.mass);
        append(d.positions_x, p.rows.positions_x);
        append(d.positions_y, p.rows.positions_y);
        append(d.positions_z, p.rows.positions_z);
        append(d.velocities_x, p.rows.velocities_x);
        append(d.velocities_y, p.rows.velocities_y);
        append(d.velocities_z, p.rows.velocities_z);
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
