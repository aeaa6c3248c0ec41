// lbm_post -- the reference's post-processing programs over this build's
// output files (host only, no GPU):
//
//   lbm_post force CASENAME.force [--from T0] [--to T1]
//       ForceCalcs (bgk_channel1/StaticInlet/ForceCalcs.cpp:12-106): running
//       mean and RMS of the three force components over the record lines
//       T0 <= line < T1 of a WriteForces file, printed as
//       "FxAv = .., FxRMS = ..".
//
//   lbm_post profiles CASENAME_slice.dat --nx X --ny Y --obst-y OBY
//            --height H --diam D --umax U [--out FILE]
//       VelProf2 (bgk_channel1/StaticInlet/VelProf2.cpp:12-155): from the
//       slice file's first zone (X x Y rows, x fastest, columns X,Y,Z,u,v,w,
//       rho,velAv[0..2],ufluc,vfluc,wfluc), the cross-stream profiles of
//       <u>/U, <v>/U, u_rms/U, v_rms/U at 1, 3.5, 6 and 8.5 diameters behind
//       the obstacle (streamwise is y there) and the wake profile along
//       x = H/2, written as CASENAME.prof2 unless --out is given.
//
// Both follow the reference programs' FP32 arithmetic and output format;
// the case constants they hard-code are arguments here.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

namespace {

int usage()
{
    std::fprintf(stderr,
                 "usage: lbm_post force CASENAME.force [--from T0] [--to T1]\n"
                 "       lbm_post profiles CASENAME_slice.dat --nx X --ny Y --obst-y OBY --height H --diam D "
                 "--umax U [--out FILE]\n");
    return 2;
}

// value of "--name" in argv, or fallback
const char* opt(int argc, char** argv, const char* name, const char* fallback)
{
    for (int i = 3; i + 1 < argc; ++i)
        if (std::strcmp(argv[i], name) == 0) return argv[i + 1];
    return fallback;
}

// the comma-separated fields of one record line
std::vector<float> fields(const std::string& line)
{
    std::vector<float> out;
    std::istringstream tok(line);
    std::string t;
    while (std::getline(tok, t, ',')) out.push_back(std::strtof(t.c_str(), nullptr));
    return out;
}

int force_stats(int argc, char** argv)
{
    std::ifstream in(argv[2]);
    if (!in) {
        std::fprintf(stderr, "lbm_post: cannot read %s\n", argv[2]);
        return 1;
    }
    std::vector<std::array<float, 3>> F;
    std::string line;
    while (std::getline(in, line)) {
        const std::vector<float> v = fields(line);
        if (v.size() < 4) continue;
        F.push_back({v[1], v[2], v[3]});
    }
    const long long n = (long long)F.size();
    const long long t0 = std::atoll(opt(argc, argv, "--from", "0"));
    const long long to = std::atoll(opt(argc, argv, "--to", "-1"));
    const long long t1 = to < 0 ? n : std::min(to, n);
    if (t0 < 0 || t0 >= t1) {
        std::fprintf(stderr, "lbm_post: empty window [%lld, %lld) of %zu lines\n", t0, t1, F.size());
        return 1;
    }
    // running forms of ForceCalcs.cpp:87-99 (FP32, j = line - T0)
    float av[3] = {0, 0, 0}, fl[3] = {0, 0, 0};
    for (long long i = t0; i < t1; ++i) {
        const int j = (int)(i - t0);
        for (int a = 0; a < 3; ++a) av[a] = av[a] * j / (j + 1) + F[i][a] / (j + 1);
    }
    for (long long i = t0; i < t1; ++i) {
        const int j = (int)(i - t0);
        for (int a = 0; a < 3; ++a) fl[a] = fl[a] * j / (j + 1) + (F[i][a] - av[a]) * (F[i][a] - av[a]) / (j + 1);
    }
    const char* name[3] = {"Fx", "Fy", "Fz"};
    for (int a = 0; a < 3; ++a) std::cout << name[a] << "Av = " << av[a] << ", " << name[a] << "RMS = " << std::sqrt(fl[a]) << "\n";
    return 0;
}

int profiles(int argc, char** argv)
{
    const std::string slice = argv[2];
    const int X = std::atoi(opt(argc, argv, "--nx", "0")), Y = std::atoi(opt(argc, argv, "--ny", "0"));
    const float oby = std::strtof(opt(argc, argv, "--obst-y", "nan"), nullptr);
    const int H = std::atoi(opt(argc, argv, "--height", "0"));
    const float d = std::strtof(opt(argc, argv, "--diam", "0"), nullptr);
    const float umax = std::strtof(opt(argc, argv, "--umax", "0"), nullptr);
    std::string out_name = opt(argc, argv, "--out", "");
    if (out_name.empty()) {
        const std::string tail = "_slice.dat";
        out_name = (slice.size() > tail.size() && slice.compare(slice.size() - tail.size(), tail.size(), tail) == 0
                        ? slice.substr(0, slice.size() - tail.size())
                        : slice) +
                   ".prof2";
    }
    if (X <= 1 || Y <= 0 || H <= 0 || H / 2 + 1 >= X || !(d > 0) || !(umax > 0) || std::isnan(oby)) {
        std::fprintf(stderr, "lbm_post profiles: need --nx > H/2+1, --ny, --obst-y, --height, --diam, --umax > 0\n");
        return 1;
    }
    std::ifstream in(slice);
    if (!in) {
        std::fprintf(stderr, "lbm_post: cannot read %s\n", slice.c_str());
        return 1;
    }
    std::string line;
    std::getline(in, line); // VARIABLES
    std::getline(in, line); // ZONE
    std::vector<float> uav((size_t)X * Y), vav((size_t)X * Y), urms((size_t)X * Y), vrms((size_t)X * Y);
    for (int j = 0; j < Y; ++j)
        for (int i = 0; i < X; ++i) {
            if (!std::getline(in, line)) {
                std::fprintf(stderr, "lbm_post: %s ends before %d x %d rows\n", slice.c_str(), X, Y);
                return 1;
            }
            const std::vector<float> v = fields(line);
            if (v.size() < 13) {
                std::fprintf(stderr, "lbm_post: %s has no statistics columns\n", slice.c_str());
                return 1;
            }
            const size_t c = (size_t)i + (size_t)X * j;
            // VelProf2.cpp:98-101: velAv[0], velAv[1], sqrt(ufluc), sqrt(vfluc), over UMAX
            uav[c] = v[7] / umax;
            vav[c] = v[8] / umax;
            urms[c] = std::sqrt(v[10]) / umax;
            vrms[c] = std::sqrt(v[11]) / umax;
        }
    std::ofstream o(out_name);
    if (!o) {
        std::fprintf(stderr, "lbm_post: cannot write %s\n", out_name.c_str());
        return 1;
    }
    auto at = [&](int i, int j) { return (size_t)i + (size_t)X * j; };
    auto station = [&](const char* label, float k) -> bool {
        const int j = (int)(oby + d * 0.5f + 1 + k * d); // VelProf2.cpp:104-126
        if (j < 0 || j >= Y) {
            std::fprintf(stderr, "lbm_post: station %s (y = %d) outside the slice\n", label, j);
            return false;
        }
        o << label << "\n";
        for (int i = 0; i < H / 2 + 1; ++i)
            o << i / d << ", " << uav[at(i, j)] << ", " << vav[at(i, j)] << ", " << urms[at(i, j)] << ", "
              << vrms[at(i, j)] << "\n";
        return true;
    };
    if (!station("x=1d", 1.f) || !station("x=3.5d", 3.5f) || !station("x=6d", 6.f) || !station("x=8.5d", 8.5f))
        return 1;
    // wake profile along x = H/2 (VelProf2.cpp:127-136)
    o << "wake profile\n";
    const int i = H / 2;
    for (int j = (int)(oby + d * 0.5f + 1); j < oby + d * 0.5f + 1 + 8.f * d && j < Y; ++j)
        o << (j - oby - d * 0.5f) / d << ", " << uav[at(i, j)] << ", " << vav[at(i, j)] << ", " << urms[at(i, j)]
          << ", " << vrms[at(i, j)] << "\n";
    return o ? 0 : 1;
}

} // namespace

int main(int argc, char** argv)
{
    if (argc < 3) return usage();
    if (std::strcmp(argv[1], "force") == 0) return force_stats(argc, argv);
    if (std::strcmp(argv[1], "profiles") == 0) return profiles(argc, argv);
    return usage();
}
