// gen_main.cpp -- nbody_gen: writes a seeded synthetic input in the reference's
// CSV format (id,name,class,mass[kg],x,y,z[AU],vx,vy,vz[AU/d]; README.md:133-138)
// with 17 significant digits, so that the file reproduces the doubles exactly.
//
//   nbody_gen --kind=disk|plummer --n=<bodies> --seed=<u64> --out=<file.csv>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <vector>

#include "nbody_capi.h"

int main(int argc, char *argv[]) {
    std::string kind = "disk", out;
    uint64_t seed = 0x4E424F44590000ull;
    uint32_t n = 0;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        const size_t eq = a.find('=');
        if (a.rfind("--", 0) != 0 || eq == std::string::npos) {
            std::fprintf(stderr, "usage: nbody_gen --kind=disk|plummer --n=N --seed=S --out=F\n");
            return 2;
        }
        const std::string k = a.substr(2, eq - 2), v = a.substr(eq + 1);
        if (k == "kind") kind = v;
        else if (k == "n") n = (uint32_t)std::stoul(v);
        else if (k == "seed") seed = std::stoull(v, nullptr, 0);
        else if (k == "out") out = v;
        else {
            std::fprintf(stderr, "unknown option --%s\n", k.c_str());
            return 2;
        }
    }
    if (n == 0 || out.empty() || (kind != "disk" && kind != "plummer")) {
        std::fprintf(stderr, "usage: nbody_gen --kind=disk|plummer --n=N --seed=S --out=F\n");
        return 2;
    }
    std::vector<double> x(n), y(n), z(n), vx(n), vy(n), vz(n), m(n);
    const int rc = kind == "disk"
                       ? nb_gen_star_disk(n, seed, x.data(), y.data(), z.data(), vx.data(),
                                          vy.data(), vz.data(), m.data())
                       : nb_gen_plummer(n, seed, x.data(), y.data(), z.data(), vx.data(),
                                        vy.data(), vz.data(), m.data());
    if (rc != NB_OK) {
        std::fprintf(stderr, "generator failed: %s\n", nb_strerror(rc));
        return 1;
    }
    FILE *f = std::fopen(out.c_str(), "w");
    if (!f) {
        std::fprintf(stderr, "cannot write %s\n", out.c_str());
        return 1;
    }
    std::fprintf(f, "id,name,class,mass,x,y,z,vx,vy,vz\n");
    for (uint32_t i = 0; i < n; ++i) {
        const bool star = kind == "disk" && i == 0;
        const char *name = kind == "disk" ? (star ? "Sun" : "body") : "star";
        const char *cls = kind == "disk" ? (star ? "STA" : "AST") : "STA";
        std::fprintf(f, "%u,%s%s,%s,%.17g,%.17g,%.17g,%.17g,%.17g,%.17g,%.17g\n", i, name,
                     star ? "" : std::to_string(i).c_str(), cls, m[i], x[i], y[i], z[i], vx[i],
                     vy[i], vz[i]);
    }
    std::fclose(f);
    return 0;
}
