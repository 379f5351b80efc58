// nBodyAlgorithm.hpp -- algorithm base class of the drop-in application.
//
// Public surface of the reference's src/simulationBackend/nBodyAlgorithm.hpp:
// 19-187 (members, constructor, startSimulation, ParaView/CSV output, energy,
// stored time series).  The SYCL queue and buffers of the reference are
// replaced by one libnbody_hip.so context (include/nbody_capi.h) that keeps
// the whole state in HBM; the host only copies data out on visualised steps.
#ifndef NBODY_HOST_NBODYALGORITHM_HPP
#define NBODY_HOST_NBODYALGORITHM_HPP

#include <array>
#include <fstream>
#include <map>
#include <string>
#include <vector>

#include "Configuration.hpp"
#include "ParaViewWriter.hpp"
#include "SimulationData.hpp"
#include "TimeMeasurement.hpp"
#include "nbody_capi.h"

class nBodyAlgorithm {
   public:
    std::string description;
    std::string outputDirectory;
    double dt;
    double t_end;
    double visualizationStepWidth;
    double G;

    TimeMeasurement timer;

    // visualised step -> per-body values
    std::map<d_type::int_t, std::vector<double>> positions_x;
    std::map<d_type::int_t, std::vector<double>> positions_y;
    std::map<d_type::int_t, std::vector<double>> positions_z;
    std::map<d_type::int_t, std::vector<double>> velocities_x;
    std::map<d_type::int_t, std::vector<double>> velocities_y;
    std::map<d_type::int_t, std::vector<double>> velocities_z;
    std::map<d_type::int_t, std::vector<double>> acceleration;  // |a|

    std::map<d_type::int_t, double> kineticEnergy;
    std::map<d_type::int_t, double> potentialEnergy;
    std::map<d_type::int_t, double> totalEnergy;
    std::map<d_type::int_t, double> virialEquilibrium;

    // G = 6.67428e-11 m^3 kg^-1 s^-2 in AU^3 kg^-1 d^-2 (nBodyAlgorithm.hpp:55-61)
    nBodyAlgorithm(double dt, double t_end, double visualizationStepWidth,
                   std::string &outputDirectory);
    virtual ~nBodyAlgorithm();

    virtual void startSimulation(const SimulationData &simulationData) = 0;

    // times.json, lastState.csv, simulation.pvd and one .vtp per stored step
    // (nBodyAlgorithm.cpp:129-251, 286-411)
    void generateParaViewOutput(const SimulationData &simulationData);
    void outputLastState(const std::string &path);

    // energies of the current device state, stored for `currentStep`
    void computeEnergy(nb_ctx *ctx, d_type::int_t currentStep);
    // |a| per body, stored for `currentStep` (nBodyAlgorithm.cpp:88-102)
    void storeAccelerations(d_type::int_t currentStep, const std::vector<double> &ax,
                            const std::vector<double> &ay, const std::vector<double> &az);
    void storeAccelerations(d_type::int_t currentStep, size_t n, const double *ax,
                            const double *ay, const double *az);
    // stores positions, velocities and |a| of visualised step `step` from the
    // nine arrays of a snapshot (x, y, z, vx, vy, vz, ax, ay, az)
    void storeSnapshot(d_type::int_t step, const std::array<const double *, 9> &a);
    // subtracts the centre-of-mass velocity from the stored step-0 velocities
    // only (nBodyAlgorithm.cpp:104-127)
    void adjustVelocities(const SimulationData &simulationData);
    // views of the stored steps in step order, for the output writers
    std::vector<paraview::StoredStep> storedSteps() const;

   protected:
    // The reference's time loop (NaiveAlgorithm.cpp:15-260,
    // BarnesHutAlgorithm.cpp:18-278); the two algorithms differ only in
    // their force evaluation, selected by `algorithm` (NB_ALGO_*).
    void runSimulation(const SimulationData &simulationData, int algorithm);

   private:
    void recordStepTimes(nb_ctx *ctx, int algorithm, bool loop);
    // the octree builder's sequences from phase times ms[NB_PHASE_COUNT]
    void recordBuildPhases(const double *ms);
};

// Throws std::runtime_error carrying the library's message when rc != NB_OK.
void nbCheck(nb_ctx *ctx, int rc, const char *what);

#endif
