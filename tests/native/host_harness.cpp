// Host-only harness of the C++ facade's parser, writers, squirmer helpers and
// analytic solutions (lbm_solver_3d_amd/host: params.cpp, io.cpp,
// squirmer.cpp, analytic.cpp), built by the CPU tests without the device
// library -- once plainly and once with AddressSanitizer and
// UndefinedBehaviorSanitizer.
//
//   host_harness decks <out_dir> <deck.in>...   parse, override, validate, print
//   host_harness bad <out_dir>                  malformed decks must throw
//   host_harness writers <out_dir>              VTK, trajectory, state files
//   host_harness poiseuille <nx> <ny> <nz> <wy> <wz> <nu> <G> <flag.bin> <ux.bin> <base>
//                                               L2 error + profiles of a field
//   host_harness duct <a> <b> <G> <mu> <y> <z>...  series values
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "analytic.hpp"
#include "io.hpp"
#include "params.hpp"
#include "squirmer.hpp"

using namespace lbm;

namespace {

int decks(int argc, char** argv) {
    const std::string out = argv[2];
    for (int k = 3; k < argc; ++k) {
        Params p;
        p.loadFile(argv[k]);
        // command-line overrides after the file, as the driver applies them
        std::string o1 = "steps=7", o2 = "output_dir=" + out, o3 = "precision=fp16";
        char* over[] = {argv[0], o1.data(), o2.data(), o3.data()};
        p.parseCommandLine(4, over);
        if (p.steps != 7 || p.outputDir != out || p.precision != 2) throw std::runtime_error("override lost");
        p.validate();
        std::FILE* devnull = std::fopen("/dev/null", "w");
        p.print(devnull);
        std::fclose(devnull);
        std::printf("%s %d %d %d %d\n", argv[k], p.NX, p.NY, p.NZ, (int)p.squirmer);
    }
    return 0;
}

int bad(int, char** argv) {
    const std::string dir = argv[2];
    const char* cases[] = {
        "nx = 4\nwalls_y = maybe\n",      // boolean
        "nx = 4\nunknown_key = 1\n",      // key
        "nx = 4\nny\n",                   // no '='
        "nx = four\n",                    // number
        "tau = 0.5\n",                    // validate: tau
        "squirmer = 1\nradius = 1\n",     // validate: radius
        "plane = yz\n",                   // extension key value
        "precision = fp8\n",              // extension key value
    };
    int thrown = 0, k = 0;
    for (const char* text : cases) {
        const std::string path = dir + "/bad" + std::to_string(k++) + ".in";
        std::ofstream(path) << text;
        try {
            Params p;
            p.loadFile(path);
            p.validate();
        } catch (const std::runtime_error& e) {
            ++thrown;
            std::printf("error: %s\n", e.what());
        }
    }
    try {
        Params p;
        p.loadFile(dir + "/does_not_exist.in");
    } catch (const std::runtime_error& e) {
        ++thrown;
        std::printf("error: %s\n", e.what());
    }
    return thrown == (int)(sizeof cases / sizeof cases[0]) + 1 ? 0 : 1;
}

int writers(int, char** argv) {
    const std::string dir = std::string(argv[2]) + "/nested/out";
    ensureDirectory(dir);
    const int NX = 5, NY = 4, NZ = 3;
    const size_t N = (size_t)NX * NY * NZ;
    std::vector<uint8_t> flag(N);
    std::vector<double> rho(N), ux(N), uy(N), uz(N);
    for (size_t n = 0; n < N; ++n) {
        flag[n] = (uint8_t)(n % 7 == 0 ? WALL : (n % 5 == 0 ? PARTICLE : FLUID));
        rho[n] = 1.0 + 1e-3 * (double)n;
        ux[n] = 1e-4 * (double)n;
        uy[n] = -2e-4 * (double)n;
        uz[n] = 3e-4;
    }
    writeVTKFields(dir + "/f.vtk", NX, NY, NZ, flag.data(), rho.data(), ux.data(), uy.data(), uz.data());

    Params p;
    p.squirmer = true;
    p.R = 4.5;
    p.beta = -2.0;
    p.planeXZ = true;
    p.angle = 20.0;
    Squirmer s;
    s.init(p, 1.0 / 6.0, 10.25, 8.5, 7.75);
    double r[3] = {3.0, -1.0, 2.0}, ub[3];
    s.surfaceVelocity(r, ub, true);
    s.Fh[0] = 1e-4; s.Th[2] = -2e-4; s.Fw[2] = softWallForce(0.3, s.eps, s.FS, s.hc);
    {
        TrajectoryWriter tw(dir + "/traj.dat");
        for (int t = 50; t <= 150; t += 50) {
            s.integrate();
            tw.write(t, s);
        }
        tw.flush();
    }
    writeSquirmerStateOf(s, 1.0 / 6.0, 150, dir + "/state.dat");
    std::printf("%.17g %.17g %.17g\n", ub[0], ub[1], ub[2]);
    std::printf("%.17g %.17g %.17g\n", s.X[0], s.X[1], s.X[2]);
    return 0;
}

std::vector<char> slurp(const char* path) {
    std::ifstream in(path, std::ios::binary);
    return std::vector<char>((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
}

int poiseuille(int, char** argv) {
    const int NX = std::atoi(argv[2]), NY = std::atoi(argv[3]), NZ = std::atoi(argv[4]);
    const bool wy = std::atoi(argv[5]) != 0, wz = std::atoi(argv[6]) != 0;
    const double nu = std::strtod(argv[7], nullptr), G = std::strtod(argv[8], nullptr);
    const std::vector<char> fl = slurp(argv[9]), uxb = slurp(argv[10]);
    const size_t N = (size_t)NX * NY * NZ;
    if (fl.size() != N || uxb.size() != N * sizeof(double)) throw std::runtime_error("field sizes");
    std::vector<double> ux(N);
    std::copy(uxb.begin(), uxb.end(), reinterpret_cast<char*>(ux.data()));
    const double e = poiseuilleErrorOf(NX, NY, NZ, wy, wz, nu, G, reinterpret_cast<const uint8_t*>(fl.data()),
                                       ux.data(), argv[11]);
    std::printf("%.12e\n", e);
    return 0;
}

int duct(int argc, char** argv) {
    const double a = std::strtod(argv[2], nullptr), b = std::strtod(argv[3], nullptr);
    const double G = std::strtod(argv[4], nullptr), mu = std::strtod(argv[5], nullptr);
    for (int k = 6; k + 1 < argc; k += 2)
        std::printf("%.17g\n", ductVelocity(std::strtod(argv[k], nullptr), std::strtod(argv[k + 1], nullptr), a, b, G, mu));
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 3) return 2;
    const std::string mode = argv[1];
    try {
        if (mode == "decks") return decks(argc, argv);
        if (mode == "bad") return bad(argc, argv);
        if (mode == "writers") return writers(argc, argv);
        if (mode == "poiseuille" && argc == 12) return poiseuille(argc, argv);
        if (mode == "duct") return duct(argc, argv);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "unexpected error: %s\n", e.what());
        return 3;
    }
    return 2;
}
