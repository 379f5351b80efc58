// ParaViewWriter.hpp -- the output files of a simulation: simulation.pvd, one
// simulation_step<i>.vtp per stored step and lastState.csv, in the layout of
// the reference's nBodyAlgorithm.cpp:129-251 and 286-411 (ASCII VTK PolyData:
// points, body ids, velocities, |a|, masses, names as character codes, orbit
// classes, vertex cells and the four energies as FieldData).
//
// Numbers are printed as a std::ostream with default flags prints them:
// doubles as printf's "%.6g" (std::to_chars, general format, precision 6, is
// defined by the C++ standard as that conversion), integers in decimal.  The
// bodies of a file are cut into chunks that worker threads format into
// private buffers, which are then written in order: the bytes do not depend on
// the thread count or the chunk size.  Small files (one chunk) are formatted
// one per thread instead.
#ifndef NBODY_HOST_PARAVIEWWRITER_HPP
#define NBODY_HOST_PARAVIEWWRITER_HPP

#include <cstddef>
#include <string>
#include <vector>

#include "SimulationData.hpp"

namespace paraview {

// One stored step: per-body arrays of n values and the FieldData energies
// (zero when energies are not computed).
struct StoredStep {
    size_t n = 0;
    const double *px = nullptr, *py = nullptr, *pz = nullptr;
    const double *vx = nullptr, *vy = nullptr, *vz = nullptr;
    const double *acc = nullptr;  // |a| per body
    double kinetic = 0, potential = 0, total = 0, virial = 0;
};

// Bodies per formatting chunk (a multiple of nothing in particular: any value
// gives the same bytes).
constexpr size_t kChunkBodies = 32768;

// The text of simulation_step<i>.vtp for `s`; masses, names and classes come
// from `d` (nBodyAlgorithm.cpp:158-247, 286-395).
std::string formatVtp(const StoredStep &s, const SimulationData &d, unsigned threads,
                      size_t chunk = kChunkBodies);

// The text of lastState.csv (nBodyAlgorithm.cpp:397-411).
std::string formatLastState(const StoredStep &s, unsigned threads, size_t chunk = kChunkBodies);

// The text of simulation.pvd for `count` stored steps (nBodyAlgorithm.cpp:148-166, 248).
std::string formatPvd(size_t count);

// Writes simulation.pvd, simulation_step<i>.vtp for every step and
// lastState.csv (from the last step) into `dir`, which must exist.
void writeSeries(const std::string &dir, const std::vector<StoredStep> &steps,
                 const SimulationData &d, unsigned threads, size_t chunk = kChunkBodies);

}  // namespace paraview

#endif
