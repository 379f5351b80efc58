// main.cpp -- N_Body_Simulation command line (same options as the reference's
// src/main.cpp:13-97 plus --precision), dispatching to NaiveAlgorithm or
// BarnesHutAlgorithm (main.cpp:226-238).
//
//   N_Body_Simulation --file=<csv> --dt=1h --t_end=365d --vs=1d --vs_dir=<dir>
//                     --algorithm=naive|BarnesHut [options]
#include <cstdlib>
#include <iostream>
#include <map>
#include <set>
#include <stdexcept>
#include <string>

#include "BarnesHutAlgorithm.hpp"
#include "Configuration.hpp"
#include "InputParser.hpp"
#include "NaiveAlgorithm.hpp"
#include "SimulationData.hpp"
#include "TimeConverter.hpp"

namespace {

// `--name=value`, `--name value`, or a bare `--flag` for booleans.
class Options {
   public:
    Options(int argc, char *argv[], const std::set<std::string> &known,
            const std::set<std::string> &booleans) {
        for (int i = 1; i < argc; ++i) {
            std::string arg = argv[i];
            if (arg.rfind("--", 0) != 0) throw std::invalid_argument("unexpected argument " + arg);
            arg = arg.substr(2);
            std::string value;
            const size_t eq = arg.find('=');
            bool has_value = eq != std::string::npos;
            if (has_value) {
                value = arg.substr(eq + 1);
                arg = arg.substr(0, eq);
            }
            if (!known.count(arg)) throw std::invalid_argument("Option '" + arg + "' does not exist");
            if (!has_value) {
                if (i + 1 < argc && std::string(argv[i + 1]).rfind("--", 0) != 0) {
                    value = argv[++i];
                } else if (booleans.count(arg)) {
                    value = "true";
                } else {
                    throw std::invalid_argument("Option '" + arg + "' is missing an argument");
                }
            }
            values_[arg] = value;
        }
    }
    size_t count(const std::string &k) const { return values_.count(k); }
    std::string str(const std::string &k) const {
        auto it = values_.find(k);
        if (it == values_.end()) throw std::invalid_argument("Option '" + k + "' has no value");
        return it->second;
    }
    int integer(const std::string &k) const {
        const std::string v = str(k);
        size_t used = 0;
        const int r = std::stoi(v, &used);
        if (used != v.size()) throw std::invalid_argument("Option '" + k + "': not an integer");
        return r;
    }
    double real(const std::string &k) const {
        const std::string v = str(k);
        size_t used = 0;
        const double r = std::stod(v, &used);
        if (used != v.size()) throw std::invalid_argument("Option '" + k + "': not a number");
        return r;
    }
    bool boolean(const std::string &k) const {
        const std::string v = str(k);
        if (v == "true" || v == "1" || v == "yes") return true;
        if (v == "false" || v == "0" || v == "no") return false;
        throw std::invalid_argument("Option '" + k + "': expected true or false");
    }

   private:
    std::map<std::string, std::string> values_;
};

}  // namespace

int main(int argc, char *argv[]) {
    try {
        const std::set<std::string> booleans = {"energy", "sort_bodies", "use_gpus"};
        const std::set<std::string> known = {
            "file", "dt", "t_end", "vs", "vs_dir", "theta", "num_wi_octree", "num_wi_top_octree",
            "num_wi_AABB", "num_wi_com", "max_level_top_octree", "storage_size_param",
            "stack_size_param", "block_size", "algorithm", "energy", "sort_bodies", "use_gpus",
            "wg_size_barnes_hut", "opt_stage", "precision"};
        Options options(argc, argv, known, booleans);

        std::string path = options.str("file");
        std::string outputDirectoryPath = options.str("vs_dir");
        std::string dt_input = options.str("dt");
        std::string t_end_input = options.str("t_end");
        std::string vs_input = options.str("vs");
        const double dt = TimeConverter::convertToEarthDays(dt_input);
        const double t_end = TimeConverter::convertToEarthDays(t_end_input);
        const double visualizationStepWidth = TimeConverter::convertToEarthDays(vs_input);

        SimulationData simulationData;
        InputParser::parse_input(path, simulationData);
        if (simulationData.mass.empty()) throw std::invalid_argument("input file holds no bodies");

        const int storageSizeParam =
            options.count("storage_size_param") ? options.integer("storage_size_param") : 16;
        const int stackSizeParam =
            options.count("stack_size_param") ? options.integer("stack_size_param") : 16;
        if (storageSizeParam < 1 || stackSizeParam < 1)
            throw std::invalid_argument("storage_size_param and stack_size_param must be >= 1");
        configuration::initializeConfigValues((d_type::int_t)simulationData.mass.size(),
                                              storageSizeParam, stackSizeParam);
        if (options.count("energy")) configuration::setEnergyComputation(options.boolean("energy"));
        if (options.count("use_gpus")) configuration::setDeviceGPU(options.boolean("use_gpus"));
        if (options.count("precision")) {
            const std::string p = options.str("precision");
            if (p == "fp64") configuration::setPrecision(NB_PREC_FP64);
            else if (p == "fp32") configuration::setPrecision(NB_PREC_FP32C);
            else throw std::invalid_argument("precision must be <fp64> or <fp32>");
        }

        const std::string algorithm = options.str("algorithm");
        if (algorithm == "naive") {
            if (options.count("block_size")) configuration::setBlockSize(options.integer("block_size"));
            if (options.count("opt_stage")) {
                const int stage = options.integer("opt_stage");
                if (stage > 2 || stage < 0)
                    throw std::invalid_argument("Optimization stage must be 0,1 or 2");
                configuration::setOptimizationStage(stage);
            }
            std::cout << "Naive algorithm configuration:" << std::endl;
            std::cout << "Block Size ------------------------------------ "
                      << configuration::naive_algorithm::blockSize << std::endl;
            std::cout << "Optimization stage acceleration kernel -------- "
                      << configuration::naive_algorithm::optimization_stage << std::endl;
            std::cout << std::endl << std::endl;
        } else if (algorithm == "BarnesHut") {
            if (options.count("theta")) configuration::setTheta(options.real("theta"));
            if (options.count("num_wi_octree"))
                configuration::setOctreeWorkItemCount(options.integer("num_wi_octree"));
            if (options.count("num_wi_top_octree"))
                configuration::setOctreeTopWorkItemCount(options.integer("num_wi_top_octree"));
            if (options.count("num_wi_com"))
                configuration::setCenterOfMassWorkItemCount(options.integer("num_wi_com"));
            if (options.count("max_level_top_octree"))
                configuration::setMaxBuildLevel(options.integer("max_level_top_octree"));
            if (options.count("num_wi_AABB"))
                configuration::setAABBWorkItemCount(options.integer("num_wi_AABB"));
            if (options.count("sort_bodies"))
                configuration::setSortBodies(options.boolean("sort_bodies"));
            if (options.count("wg_size_barnes_hut"))
                configuration::setWorkGroupSizeBarnesHut(options.integer("wg_size_barnes_hut"));
            namespace bh = configuration::barnes_hut_algorithm;
            std::cout << "Barnes-Hut algorithm configuration:" << std::endl;
            std::cout << "Theta ----------------------------------- " << bh::theta << std::endl;
            std::cout << "Work-items AABB creation ---------------- " << bh::AABBWorkItemCount << std::endl;
            std::cout << "Work-items octree creation -------------- " << bh::octreeWorkItemCount << std::endl;
            std::cout << "Work-items center of mass calculation --- " << bh::centerOfMassWorkItemCount
                      << std::endl;
            std::cout << "Work-items top of octree creation ------- " << bh::octreeTopWorkItemCount
                      << std::endl;
            std::cout << "Maximum build level top of octree ------- " << bh::maxBuildLevel << std::endl;
            std::cout << "Work-group size acceleration kernel ----- " << bh::workGroupSize << std::endl;
            std::cout << "Sort Bodies enabled --------------------- " << bh::sortBodies << std::endl;
            std::cout << std::endl << std::endl;
        } else {
            throw std::invalid_argument("Algorithm must either be <naive> or <BarnesHut>");
        }
        std::cout << "Precision ------------------------------- "
                  << (configuration::precision == NB_PREC_FP64 ? "fp64" : "fp32 (fp64 accumulation)")
                  << std::endl << std::endl;

        if (algorithm == "naive") {
            NaiveAlgorithm a(dt, t_end, visualizationStepWidth, outputDirectoryPath);
            a.startSimulation(simulationData);
            a.generateParaViewOutput(simulationData);
        } else {
            BarnesHutAlgorithm a(dt, t_end, visualizationStepWidth, outputDirectoryPath);
            a.startSimulation(simulationData);
            a.generateParaViewOutput(simulationData);
        }
    } catch (const std::exception &e) {
        std::cerr << "error: " << e.what() << std::endl;
        return 1;
    }
    return 0;
}
