#include "nBodyAlgorithm.hpp"

#include <algorithm>
#include <array>
#include <exception>
#include <thread>
#include <cmath>
#include <ctime>
#include <filesystem>
#include <iostream>
#include <stdexcept>
#include <utility>

#include "Parallel.hpp"

void nbCheck(nb_ctx *ctx, int rc, const char *what) {
    if (rc == NB_OK) return;
    std::string msg = std::string(what) + " failed: " + nb_strerror(rc);
    if (ctx) msg += std::string(" (") + nb_last_error(ctx) + ")";
    throw std::runtime_error(msg);
}

nBodyAlgorithm::nBodyAlgorithm(double dt, double t_end, double visualizationStepWidth,
                               std::string &outputDirectory)
    : outputDirectory(outputDirectory), dt(dt), t_end(t_end),
      visualizationStepWidth(visualizationStepWidth) {
    G = 6.67428 * std::pow(10, -11);
    const double meter_AU = 1.0 / (1.49597870691 * std::pow(10, 11));
    const double second_Days = 1.0 / 86400;
    G = G * (std::pow(meter_AU, 3) / std::pow(second_Days, 2));
}

nBodyAlgorithm::~nBodyAlgorithm() = default;

namespace {

double phase(nb_ctx *ctx, int p) {
    double ms[NB_PHASE_COUNT];
    nbCheck(ctx, nb_phase_ms(ctx, ms), "nb_phase_ms");
    return ms[p];
}

// One host thread that stores a snapshot; joined before the next one starts
// and, at the latest, before the context (and its pinned slots) is destroyed.
class SnapshotCopier {
   public:
    template <class F>
    void start(F f) {
        thread_ = std::thread([this, f] {
            try {
                f();
            } catch (...) {
                error_ = std::current_exception();
            }
        });
    }
    void join() {
        if (thread_.joinable()) thread_.join();
        if (error_) {
            std::exception_ptr e = error_;
            error_ = nullptr;
            std::rethrow_exception(e);
        }
    }
    ~SnapshotCopier() {
        if (thread_.joinable()) thread_.join();
    }

   private:
    std::thread thread_;
    std::exception_ptr error_;
};

struct CtxGuard {
    nb_ctx *ctx = nullptr;
    ~CtxGuard() {
        if (ctx) nb_destroy(ctx);
    }
};

}  // namespace

void nBodyAlgorithm::recordStepTimes(nb_ctx *ctx, int algorithm, bool loop) {
    double ms[NB_PHASE_COUNT];
    nbCheck(ctx, nb_phase_ms(ctx, ms), "nb_phase_ms");
    if (loop) {
        timer.addTimeToSequence("Leapfrog Part 1", ms[NB_PHASE_LF1]);
        timer.addTimeToSequence("Leapfrog Part 2", ms[NB_PHASE_LF2]);
    }
    if (algorithm == NB_ALGO_BARNES_HUT) {
        timer.addTimeToSequence("Octree creation", ms[NB_PHASE_OCTREE]);
        recordBuildPhases(ms);
        timer.addTimeToSequence("Total Time", ms[NB_PHASE_OCTREE] + ms[NB_PHASE_ACCEL]);
    }
    timer.addTimeToSequence("Acceleration Kernel Time", ms[NB_PHASE_ACCEL]);
}

// The default (Subtrees) builder's sequences, ParallelOctreeTopDownSubtrees.cpp:76-92,
// with this build's phases in their places: the octant keys descend every
// body through the top levels ("Build octree to level"); the radix sort
// places the bodies of each subtree together ("Sort bodies for subtrees");
// the cell numbering and linking build all levels below ("Build subtrees").
// No step prepares per-subtree counts, and the in-order body order is the
// radix sort's output, so "Prepare subtrees" and "Sort bodies" hold 0.
void nBodyAlgorithm::recordBuildPhases(const double *ms) {
    timer.addTimeToSequence("AABB creation", ms[NB_PHASE_AABB]);
    timer.addTimeToSequence("Build octree to level", ms[NB_PHASE_KEYS]);
    timer.addTimeToSequence("Prepare subtrees", 0.0);
    timer.addTimeToSequence("Sort bodies for subtrees", ms[NB_PHASE_SORT]);
    timer.addTimeToSequence("Build subtrees", ms[NB_PHASE_TOPOLOGY]);
    timer.addTimeToSequence("Compute center of mass", ms[NB_PHASE_COM]);
    if (configuration::barnes_hut_algorithm::sortBodies) timer.addTimeToSequence("Sort bodies", 0.0);
}

// The library's storage and stack multipliers from the reference's scaled
// sizes (the inverse of Configuration.cpp:24-33).
static std::pair<uint32_t, uint32_t> librarySizeParams(d_type::int_t N) {
    namespace bh = configuration::barnes_hut_algorithm;
    const uint64_t n = std::max<d_type::int_t>(N, 1);
    const uint64_t lg = std::max<uint64_t>(1, (uint64_t)std::ceil(std::log2((double)n)));
    const uint64_t extra = N < 15000 ? 500 : 0;
    const uint64_t storage = std::max<uint64_t>(1, (uint64_t)bh::storageSizeParameter / n);
    const uint64_t stack = bh::stackSize > extra ? std::max<uint64_t>(1, (bh::stackSize - extra) / lg) : 1;
    return {(uint32_t)std::min<uint64_t>(storage, UINT32_MAX),
            (uint32_t)std::min<uint64_t>(stack, UINT32_MAX)};
}

void nBodyAlgorithm::runSimulation(const SimulationData &simulationData, int algorithm) {
    const d_type::int_t N = configuration::numberOfBodies;
    if (!configuration::use_GPUs)
        throw std::invalid_argument(
            "use_gpus=false: this implementation runs the simulation on the GPU only");

    positions_x[0] = simulationData.positions_x;
    positions_y[0] = simulationData.positions_y;
    positions_z[0] = simulationData.positions_z;
    velocities_x[0] = simulationData.velocities_x;
    velocities_y[0] = simulationData.velocities_y;
    velocities_z[0] = simulationData.velocities_z;
    adjustVelocities(simulationData);

    nb_params p;
    nb_default_params(&p);
    p.G = G;
    p.eps2 = configuration::epsilon2;
    p.theta = configuration::barnes_hut_algorithm::theta;
    p.precision = (uint32_t)configuration::precision;
    p.sort_bodies = configuration::barnes_hut_algorithm::sortBodies ? 1 : 0;
    // The reference keeps only the scaled sizes (Configuration.cpp:24-33):
    // storageSizeParameter = storage multiplier * N, and stackSize =
    // stack multiplier * ceil(log2 N) (+ 500 below 15000 bodies); the library
    // takes the multipliers.
    const std::pair<uint32_t, uint32_t> sizes = librarySizeParams(N);
    p.storage_size_param = sizes.first;
    p.stack_size_param = sizes.second;
    CtxGuard guard;
    nbCheck(nullptr, nb_create(&guard.ctx, N, &p), "nb_create");
    nb_ctx *ctx = guard.ctx;
    nbCheck(ctx,
            nb_set_state(ctx, simulationData.positions_x.data(), simulationData.positions_y.data(),
                         simulationData.positions_z.data(), simulationData.velocities_x.data(),
                         simulationData.velocities_y.data(), simulationData.velocities_z.data(),
                         simulationData.mass.data()),
            "nb_set_state");

    char devname[256] = "unknown";
    nb_device_name(devname, sizeof devname);
    std::string device = devname;
    d_type::int_t bodyCount = N;
    timer.setProperties(description, bodyCount, device);
    timer.addTimingSequence("Acceleration Kernel Time");
    timer.addTimingSequence("Leapfrog Part 1");
    timer.addTimingSequence("Leapfrog Part 2");
    if (algorithm == NB_ALGO_BARNES_HUT) {
        // BarnesHutAlgorithm.cpp:81-100 with the default builder
        for (const char *key : {"Total Time", "Octree creation", "AABB creation",
                                "Compute center of mass", "Build octree to level",
                                "Prepare subtrees", "Sort bodies for subtrees", "Build subtrees"})
            timer.addTimingSequence(key);
        if (configuration::barnes_hut_algorithm::sortBodies) timer.addTimingSequence("Sort bodies");
    }

    std::vector<double> ax(N), ay(N), az(N);
    double time = 0.0;
    double timeSinceLastVisualization = 0.0;
    d_type::int_t currentStep = 0;

    // initial accelerations (NaiveAlgorithm.cpp:86-113, BarnesHutAlgorithm.cpp:105-133)
    if (algorithm == NB_ALGO_BARNES_HUT) {
        nbCheck(ctx, nb_bh_build(ctx), "nb_bh_build");
        double build_ms[NB_PHASE_COUNT];
        nbCheck(ctx, nb_phase_ms(ctx, build_ms), "nb_phase_ms");
        const double octree = build_ms[NB_PHASE_OCTREE];
        timer.addTimeToSequence("Octree creation", octree);
        recordBuildPhases(build_ms);
        nbCheck(ctx, nb_bh_forces(ctx), "nb_bh_forces");
        const double accel = phase(ctx, NB_PHASE_ACCEL);
        timer.addTimeToSequence("Acceleration Kernel Time", accel);
        timer.addTimeToSequence("Total Time", octree + accel);
        std::cout << std::endl;
    } else {
        nbCheck(ctx, nb_forces_direct(ctx), "nb_forces_direct");
        timer.addTimeToSequence("Acceleration Kernel Time", phase(ctx, NB_PHASE_ACCEL));
    }
    if (configuration::compute_energy) computeEnergy(ctx, currentStep);
    nbCheck(ctx, nb_get_accel(ctx, ax.data(), ay.data(), az.data()), "nb_get_accel");
    storeAccelerations(currentStep, ax, ay, az);
    std::cout << "Finished initial step " << currentStep << std::endl << std::endl;

    time += dt;
    timeSinceLastVisualization += dt;
    currentStep += 1;

    // A visualised step's state is copied out asynchronously: the DMA runs
    // while the GPU computes the next step, after which a host thread stores
    // it into the time series during the step after that.  The library keeps
    // two pinned slots, so the thread must be done before the second begin
    // after the snapshot it reads.
    SnapshotCopier copier;
    d_type::int_t pendingSnapshot = 0;
    bool snapshotPending = false;
    auto takeSnapshot = [&] {
        std::array<const double *, 9> a;
        nbCheck(ctx, nb_snapshot_wait(ctx, a.data()), "nb_snapshot_wait");
        copier.join();
        const d_type::int_t step = pendingSnapshot;
        copier.start([this, step, a] { storeSnapshot(step, a); });
        snapshotPending = false;
    };
    while (time <= t_end + 0.000001) {
        const bool visualizeCurrentStep =
            (std::abs(timeSinceLastVisualization - visualizationStepWidth) < 0.000001);
        // kick-drift, [octree], forces, kick -- one synchronisation per step
        nbCheck(ctx, nb_steps(ctx, algorithm, dt, 1), "nb_steps");
        recordStepTimes(ctx, algorithm, true);
        if (snapshotPending) takeSnapshot();
        if (visualizeCurrentStep) {
            std::cout << "Finished step " << currentStep << std::endl << std::endl;
            nbCheck(ctx, nb_snapshot_begin(ctx), "nb_snapshot_begin");
            pendingSnapshot = currentStep;
            snapshotPending = true;
            if (configuration::compute_energy) computeEnergy(ctx, currentStep);
            currentStep += 1;
            timeSinceLastVisualization = 0.0;
        }
        time += dt;
        timeSinceLastVisualization += dt;
    }
    if (snapshotPending) takeSnapshot();
    copier.join();
}

void nBodyAlgorithm::computeEnergy(nb_ctx *ctx, d_type::int_t currentStep) {
    double k, u, e, v;
    nbCheck(ctx, nb_energy(ctx, &k, &u, &e, &v), "nb_energy");
    kineticEnergy[currentStep] = k;
    potentialEnergy[currentStep] = u;
    totalEnergy[currentStep] = e;
    virialEquilibrium[currentStep] = v;
}

void nBodyAlgorithm::storeAccelerations(d_type::int_t currentStep, const std::vector<double> &ax,
                                        const std::vector<double> &ay,
                                        const std::vector<double> &az) {
    storeAccelerations(currentStep, ax.size(), ax.data(), ay.data(), az.data());
}

void nBodyAlgorithm::storeAccelerations(d_type::int_t currentStep, size_t n, const double *ax,
                                        const double *ay, const double *az) {
    std::vector<double> &out = acceleration[currentStep];
    out.resize(n);
    for (size_t i = 0; i < n; ++i) out[i] = std::sqrt(ax[i] * ax[i] + ay[i] * ay[i] + az[i] * az[i]);
}

void nBodyAlgorithm::storeSnapshot(d_type::int_t step, const std::array<const double *, 9> &a) {
    const size_t n = configuration::numberOfBodies;
    std::vector<double> *dst[6] = {&positions_x[step],  &positions_y[step],  &positions_z[step],
                                   &velocities_x[step], &velocities_y[step], &velocities_z[step]};
    std::vector<double> &acc = acceleration[step];
    host_parallel::parallelFor(7, host_parallel::defaultThreads(), [&](size_t k) {
        if (k < 6) {
            dst[k]->assign(a[k], a[k] + n);
        } else {
            acc.resize(n);
            for (size_t i = 0; i < n; ++i)
                acc[i] = std::sqrt(a[6][i] * a[6][i] + a[7][i] * a[7][i] + a[8][i] * a[8][i]);
        }
    });
}

void nBodyAlgorithm::adjustVelocities(const SimulationData &simulationData) {
    double sumMasses = 0, smx = 0, smy = 0, smz = 0;
    for (d_type::int_t i = 0; i < configuration::numberOfBodies; ++i) {
        sumMasses += simulationData.mass[i];
        smx += simulationData.mass[i] * simulationData.velocities_x[i];
        smy += simulationData.mass[i] * simulationData.velocities_y[i];
        smz += simulationData.mass[i] * simulationData.velocities_z[i];
    }
    const double ui_x = smx / sumMasses, ui_y = smy / sumMasses, ui_z = smz / sumMasses;
    for (d_type::int_t i = 0; i < configuration::numberOfBodies; ++i) {
        velocities_x[0][i] -= ui_x;
        velocities_y[0][i] -= ui_y;
        velocities_z[0][i] -= ui_z;
    }
}

std::vector<paraview::StoredStep> nBodyAlgorithm::storedSteps() const {
    std::vector<paraview::StoredStep> steps;
    auto energy = [](const std::map<d_type::int_t, double> &values, d_type::int_t i) {
        if (!configuration::compute_energy) return 0.0;
        auto it = values.find(i);
        return it != values.end() ? it->second : 0.0;
    };
    for (d_type::int_t i = 0; i < (d_type::int_t)positions_x.size(); ++i) {
        paraview::StoredStep s;
        s.n = positions_x.at(i).size();
        s.px = positions_x.at(i).data();
        s.py = positions_y.at(i).data();
        s.pz = positions_z.at(i).data();
        s.vx = velocities_x.at(i).data();
        s.vy = velocities_y.at(i).data();
        s.vz = velocities_z.at(i).data();
        s.acc = acceleration.at(i).data();
        s.kinetic = energy(kineticEnergy, i);
        s.potential = energy(potentialEnergy, i);
        s.total = energy(totalEnergy, i);
        s.virial = energy(virialEquilibrium, i);
        steps.push_back(s);
    }
    return steps;
}

void nBodyAlgorithm::generateParaViewOutput(const SimulationData &simulationData) {
    std::time_t currentTime = std::time(nullptr);
    std::string time = static_cast<std::string>(std::ctime(&currentTime));
    std::replace(time.begin(), time.end(), ' ', '_');
    time = time.substr(0, time.size() - 1);
    const std::string filePathBase = outputDirectory + '/' + time + '/';
    std::filesystem::create_directories(filePathBase);

    timer.exportJSON(filePathBase + "times.json");
    // lastState.csv, simulation.pvd and one .vtp per stored step, formatted
    // on the host's worker threads
    paraview::writeSeries(filePathBase, storedSteps(), simulationData,
                          host_parallel::defaultThreads());
}

void nBodyAlgorithm::outputLastState(const std::string &path) {
    const std::vector<paraview::StoredStep> steps = storedSteps();
    if (steps.empty()) throw std::runtime_error("outputLastState: no stored step");
    std::ofstream csvFile(path, std::ios::binary);
    csvFile << paraview::formatLastState(steps.back(), host_parallel::defaultThreads());
    if (!csvFile) throw std::runtime_error("write to " + path + " failed");
}
