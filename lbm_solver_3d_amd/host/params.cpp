// params.cpp -- input parsing and validation with the reference's semantics
// (src/utils/params.cpp): case-insensitive keys with the same aliases,
// booleans 1/0 true/false yes/no on/off, '#' comments, "key = value" lines,
// command-line key=value overrides applied in order, and the same
// validation rules and messages.
#include "params.hpp"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <fstream>
#include <stdexcept>

namespace lbm {

namespace {

std::string strip(const std::string& s) {
    const char* ws = " \t\r\n";
    const size_t b = s.find_first_not_of(ws);
    if (b == std::string::npos) return std::string();
    return s.substr(b, s.find_last_not_of(ws) - b + 1);
}

std::string lowercase(std::string s) {
    for (char& ch : s) ch = (char)std::tolower((unsigned char)ch);
    return s;
}

bool parse_bool(const std::string& v) {
    const std::string l = lowercase(v);
    if (l == "1" || l == "true" || l == "yes" || l == "on") return true;
    if (l == "0" || l == "false" || l == "no" || l == "off") return false;
    throw std::runtime_error("cannot interpret '" + v + "' as a boolean");
}

int parse_precision(const std::string& v) {
    const std::string l = lowercase(v);
    if (l == "fp64" || l == "0") return 0;
    if (l == "fp32" || l == "1") return 1;
    if (l == "fp16" || l == "2") return 2;
    if (l == "bf16" || l == "3") return 3;
    throw std::runtime_error("precision must be fp64, fp32, fp16 or bf16");
}

}  // namespace

void Params::set(const std::string& keyIn, const std::string& v) {
    const std::string key = lowercase(strip(keyIn));
    auto I = [&] { return std::stoi(v); };
    auto D = [&] { return std::stod(v); };
    try {
        if (key == "nx") NX = I();
        else if (key == "ny") NY = I();
        else if (key == "nz") NZ = I();
        else if (key == "tau") tau = D();
        else if (key == "lambda") Lambda = D();
        else if (key == "fx") Fx = D();
        else if (key == "fy") Fy = D();
        else if (key == "fz") Fz = D();
        else if (key == "walls_y" || key == "wallsy") wallsY = parse_bool(v);
        else if (key == "walls_z" || key == "wallsz") wallsZ = parse_bool(v);
        else if (key == "steps") steps = I();
        else if (key == "print_every" || key == "print") printEvery = I();
        else if (key == "vtk_every" || key == "vtk") vtkEvery = I();
        else if (key == "traj_every" || key == "traj") trajEvery = I();
        else if (key == "tol") tol = D();
        else if (key == "mass_correction_every") massCorrectionEvery = I();
        else if (key == "output_dir") outputDir = v;
        else if (key == "prefix") prefix = v;
        else if (key == "squirmer" || key == "sq") squirmer = parse_bool(v);
        else if (key == "radius" || key == "r") R = D();
        else if (key == "b1") B1 = D();
        else if (key == "beta") beta = D();
        else if (key == "rho_p" || key == "rhop") rhoP = D();
        else if (key == "x0") x0 = D();
        else if (key == "y0") y0 = D();
        else if (key == "z0") z0 = D();
        else if (key == "angle") angle = D();
        else if (key == "planar") planar = parse_bool(v);
        else if (key == "eps") eps = D();
        else if (key == "hc") hc = D();
        else if (key == "hmin") hmin = D();
        else if (key == "precision") precision = parse_precision(v);
        else if (key == "plane") {
            const std::string l = lowercase(strip(v));
            if (l == "xy") planeXZ = false;
            else if (l == "xz") planeXZ = true;
            else throw std::runtime_error("plane must be xy or xz");
        } else if (key == "device") device = I();
        else throw std::runtime_error("unknown parameter '" + key + "'");
    } catch (const std::invalid_argument&) {
        throw std::runtime_error("invalid value '" + v + "' for parameter '" + key + "'");
    }
}

void Params::loadFile(const std::string& path) {
    std::ifstream in(path);
    if (!in) throw std::runtime_error("cannot open input file '" + path + "'");
    std::string line;
    int lineNo = 0;
    while (std::getline(in, line)) {
        ++lineNo;
        const size_t hash = line.find('#');
        if (hash != std::string::npos) line.erase(hash);
        line = strip(line);
        if (line.empty()) continue;
        const size_t eq = line.find('=');
        if (eq == std::string::npos)
            throw std::runtime_error(path + ":" + std::to_string(lineNo) + ": expected 'key = value'");
        try {
            set(line.substr(0, eq), strip(line.substr(eq + 1)));
        } catch (const std::runtime_error& e) {
            throw std::runtime_error(path + ":" + std::to_string(lineNo) + ": " + e.what());
        }
    }
}

void Params::parseCommandLine(int argc, char** argv) {
    for (int k = 1; k < argc; ++k) {
        const std::string a = argv[k];
        const size_t eq = a.find('=');
        if (eq == std::string::npos) loadFile(a);
        else set(a.substr(0, eq), a.substr(eq + 1));
    }
}

void Params::validate() const {
    auto fail = [](const char* m) { throw std::runtime_error(m); };
    if (NX < 1 || NY < 1 || NZ < 1) fail("grid dimensions must be positive");
    if (wallsY && NY < 4) fail("ny must be >= 4 with walls in y");
    if (wallsZ && NZ < 4) fail("nz must be >= 4 with walls in z");
    if (tau <= 0.5) fail("tau must be > 0.5");
    if (Lambda <= 0.0) fail("lambda must be > 0");
    if (steps < 0) fail("steps must be >= 0");
    if (printEvery < 1 || trajEvery < 1) fail("print_every and traj_every must be >= 1");
    if (!squirmer) return;
    if (R <= 1.0) fail("squirmer radius must be > 1 lattice unit");
    if (std::fabs(angle) >= 45.0) fail("initial angle must satisfy |angle| < 45 deg from the x-axis");
    if (rhoP <= 0.0) fail("rho_p must be > 0");
    if (eps < 0.0 || hc <= 0.0 || hmin < 0.0) fail("eps >= 0, hc > 0, hmin >= 0 required");
    const double y = y0 >= 0 ? y0 : 0.5 * (NY - 1);
    if (wallsY && (y - R - 0.5 < hmin || (NY - 1.5) - y - R < hmin))
        fail("squirmer initially overlaps (or is too close to) a y-wall");
    if (2.0 * R + 2.0 > NX || (!wallsY && 2.0 * R + 2.0 > NY) || 2.0 * R + 2.0 > NZ)
        fail("domain too small for the squirmer");
    if (2.0 * B1 / 3.0 > 0.05) fail("swimming speed U0 = 2 B1 / 3 too large (Mach number); keep U0 < 0.05");
}

void Params::print(std::FILE* fp) const {
    static const char* prec[] = {"fp64", "fp32", "fp16", "bf16"};
    std::fprintf(fp, "grid        : %d x %d x %d\n", NX, NY, NZ);
    std::fprintf(fp, "fluid       : tau = %.4f (nu = %.5f), TRT Lambda = %.4f, F = (%.3e, %.3e, %.3e)\n", tau,
                 (tau - 0.5) / 3.0, Lambda, Fx, Fy, Fz);
    std::fprintf(fp, "walls       : y %s, z %s (x periodic)\n", wallsY ? "on" : "off", wallsZ ? "on" : "off");
    std::fprintf(fp, "run         : %d steps, output '%s/%s_*'\n", steps, outputDir.c_str(), prefix.c_str());
    std::fprintf(fp, "device      : GPU %d, population storage %s\n", device, prec[precision & 3]);
    if (!squirmer) return;
    const double U0 = 2.0 * B1 / 3.0, nu = (tau - 0.5) / 3.0;
    std::fprintf(fp, "squirmer    : R = %.2f, B1 = %.4f, beta = %+.2f (%s), U0 = %.5f, Re = 2RU0/nu = %.3f\n", R, B1,
                 beta, beta < 0 ? "pusher" : (beta > 0 ? "puller" : "neutral"), U0, 2 * R * U0 / nu);
    std::fprintf(fp, "              angle = %.1f deg (%s plane), rho_p = %.2f, planar = %d\n", angle,
                 planeXZ ? "x-z" : "x-y", rhoP, (int)planar);
    std::fprintf(fp, "wall pot.   : eps = %.3f (F_max = eps F_S), hc = %.2f, hmin = %.2f\n", eps, hc, hmin);
}

}  // namespace lbm
