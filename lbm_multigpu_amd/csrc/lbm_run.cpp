// lbm_run -- the reference programs' main() as a host driver over liblbm.
//
// The reference configures each run with compile-time #defines and runs one
// main() per program (MRTtest/DynInlet3.cu:2834-3335): set up the lattice,
// initialise, loop TMAX steps (forces from STARTF, running statistics from
// START_VELAV), then write CASENAME.dat (WriteResults, :2601-2694),
// CASENAME.force (WriteForces, :2788-2796) and CASENAME.inputs (WriteInputs,
// :2805-2833).  This driver reads the same parameters from a `.inputs` run
// record (the WriteInputs keys, plus the "# lbm-mi355x" extension keys the
// Python layer writes) and produces the same three files through the C ABI.
//
//   lbm_run CASE.inputs [--case NAME] [--steps N] [--parse-only]
//
// --parse-only prints the parsed configuration as a `.inputs` record and
// exits without touching a GPU.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/lbm.h"

namespace {

// WriteInputs keys in the reference's order (MRTtest/DynInlet3.cu:2807-2831;
// Test/im80.inputs adds "KP:"), then the extension keys.
const char* const kRefKeys[] = {
    "Base domain size", "Base blocksize:", "Obst1 location:", "Obst1 radius:", "Obst2 location:",
    "Obst2 radius:", "RE:", "UMAX:", "omega", "DPDY", "TMAX:", "STARTF:", "START_VELAV:",
    "START_VELFLUC:", "REFINEMENT:", "MODEL:", "Smagorinsky LES:", "CS:", "LR domain size",
    "LR factor", "LR location", "LR blocksize:", "omega in LR", "GPUs per node", "KP:"};
const char* const kExtKeys[] = {
    "MRT rates:", "Boundary x:", "Boundary y:", "Boundary z:", "Halfway bounce-back:", "Storage:",
    "Geometry:", "Inlet perturbation:", "Seed:", "Temporal interpolation:", "Force steps:", "Statistics:",
    "DYNY1:", "Inlet control:", "DYNY2:", "Dynamic inlet mode:", "Trip until:", "Trip z:"};

struct Run {
    int nx = 64, ny = 64, nz = 64;
    std::string bc[3] = {"periodic", "periodic", "periodic"};
    std::string model = "BGK", rates = "lrmgpu", storage = "f32", geometry = "none";
    bool les = false, halfway = false, refinement = false, temporal = true, velav = false;
    double cs = 0, omega = 1, re = 0, umax = 0, dpdy = 0, inlet_amp = 0, lr_factor = 0.5, omega_lr = 0;
    long long tmax = 0, startf = 0, start_velav = 0, start_velfluc = 0, seed = 0;
    int force_steps = -1, gpus = 1;
    double obst1[3] = {0, 0, 0}, obst1_r = 0, obst2[3] = {0, 0, 0}, obst2_r = 0;
    int lr[3] = {0, 0, 0};
    double lr_loc[3] = {0, 0, 0};
    // dynamic-inlet precursor (MRTtest/DynInlet3.cu:1744, :2561-2600, :3147-3170)
    int dyn_y1 = -1, ctrl_every = 100, ctrl_start = 10000;
    double kp = 0;
    // the precursor form of bgk_channel1/Re3000_Cyl2 (periodic rows 0..DYNY1),
    // the im 53 ghost row DYNY2, and the trip obstacle's time window
    int dyn_mode = 0, dyn_y2 = 1, trip_until = 0, trip_z = -1;
};

std::string trim(const std::string& s)
{
    size_t a = s.find_first_not_of(" \t\r\n"), b = s.find_last_not_of(" \t\r\n");
    return a == std::string::npos ? std::string() : s.substr(a, b - a + 1);
}

// "Key <sep> value" with the reference's separators ("Key: \tv", "Key \t: v",
// "Key \tv"): the longest known key that starts the line.
bool split(const std::string& line, std::string& key, std::string& val)
{
    size_t best = 0;
    const char* hit = nullptr;
    auto consider = [&](const char* k) {
        const size_t n = std::strlen(k);
        if (line.compare(0, n, k) != 0 || n <= best) return;
        if (std::strcmp(k, "omega") == 0 && trim(line.substr(n)).rfind("in LR", 0) == 0) return;
        best = n;
        hit = k;
    };
    for (const char* k : kRefKeys) consider(k);
    for (const char* k : kExtKeys) consider(k);
    if (!hit) return false;
    key = hit;
    std::string v = trim(line.substr(best));
    if (!v.empty() && v[0] == ':') v = trim(v.substr(1));
    val = v;
    return true;
}

std::vector<double> numbers(std::string s, const char* seps)
{
    for (char& ch : s)
        if (std::strchr(seps, ch) || ch == '(' || ch == ')') ch = ' ';
    std::istringstream in(s);
    std::vector<double> v;
    double x;
    while (in >> x) v.push_back(x);
    return v;
}

bool parse(const std::string& path, Run& r)
{
    std::ifstream in(path);
    if (!in) return false;
    std::map<std::string, std::string> seen;
    std::string line;
    while (std::getline(in, line)) {
        while (!line.empty() && (line.back() == '\r' || line.back() == ' ' || line.back() == '\t')) line.pop_back();
        if (line.empty() || line[0] == '#') continue;
        std::string k, v;
        if (!split(line, k, v) || v.empty()) continue; // unknown key or a truncated line
        seen[k] = v;
    }
    auto has = [&](const char* k) { return seen.count(k) != 0; };
    auto num = [&](const char* k) { return std::atof(seen[k].c_str()); };
    if (has("Base domain size")) {
        auto v = numbers(seen["Base domain size"], "x");
        if (v.size() == 3) { r.nx = (int)v[0]; r.ny = (int)v[1]; r.nz = (int)v[2]; }
    }
    if (has("Obst1 location:")) {
        auto v = numbers(seen["Obst1 location:"], ",");
        for (size_t a = 0; a < 3 && a < v.size(); ++a) r.obst1[a] = v[a];
    }
    if (has("Obst1 radius:")) r.obst1_r = num("Obst1 radius:");
    if (has("Obst2 location:")) {
        auto v = numbers(seen["Obst2 location:"], ",");
        for (size_t a = 0; a < 3 && a < v.size(); ++a) r.obst2[a] = v[a];
    }
    if (has("Obst2 radius:")) r.obst2_r = num("Obst2 radius:");
    if (has("RE:")) r.re = num("RE:");
    if (has("UMAX:")) r.umax = num("UMAX:");
    if (has("DPDY")) r.dpdy = num("DPDY");
    if (has("TMAX:")) r.tmax = (long long)num("TMAX:");
    if (has("STARTF:")) r.startf = (long long)num("STARTF:");
    if (has("START_VELAV:")) r.start_velav = (long long)num("START_VELAV:");
    if (has("START_VELFLUC:")) r.start_velfluc = (long long)num("START_VELFLUC:");
    if (has("CS:")) r.cs = num("CS:");
    if (has("omega in LR")) r.omega_lr = num("omega in LR");
    if (has("LR factor")) r.lr_factor = num("LR factor");
    if (has("GPUs per node")) r.gpus = (int)num("GPUs per node");
    if (has("REFINEMENT:")) r.refinement = num("REFINEMENT:") != 0;
    if (has("MODEL:")) r.model = trim(seen["MODEL:"]);
    if (has("Smagorinsky LES:")) r.les = num("Smagorinsky LES:") != 0;
    if (has("LR domain size")) {
        auto v = numbers(seen["LR domain size"], "x");
        for (size_t a = 0; a < 3 && a < v.size(); ++a) r.lr[a] = (int)v[a];
    }
    if (has("LR location")) {
        auto v = numbers(seen["LR location"], "x");
        for (size_t a = 0; a < 3 && a < v.size(); ++a) r.lr_loc[a] = v[a];
    }
    if (has("omega")) r.omega = num("omega");
    else if (r.re > 0) r.omega = 1.0 / (3.0 * (r.umax * 2.0 * r.obst1_r / r.re) + 0.5); // main(), :2869
    if (has("MRT rates:")) r.rates = trim(seen["MRT rates:"]);
    if (has("Halfway bounce-back:")) r.halfway = num("Halfway bounce-back:") != 0;
    if (has("Storage:")) r.storage = trim(seen["Storage:"]);
    if (has("Geometry:")) r.geometry = trim(seen["Geometry:"]);
    if (has("Inlet perturbation:")) r.inlet_amp = num("Inlet perturbation:");
    if (has("Seed:")) r.seed = std::atoll(seen["Seed:"].c_str());
    if (has("Temporal interpolation:")) r.temporal = num("Temporal interpolation:") != 0;
    if (has("Force steps:")) r.force_steps = (int)num("Force steps:");
    if (has("Statistics:")) r.velav = num("Statistics:") != 0;
    else if (has("START_VELAV:")) r.velav = true; // programs that print START_VELAV build with VELAV 1
    const char* axis[3] = {"Boundary x:", "Boundary y:", "Boundary z:"};
    for (int a = 0; a < 3; ++a)
        if (has(axis[a])) r.bc[a] = trim(seen[axis[a]]);
    if (has("KP:")) r.kp = num("KP:");
    if (has("DYNY1:")) r.dyn_y1 = (int)num("DYNY1:");
    if (has("Inlet control:")) { // "<every>x<start>"
        auto v = numbers(seen["Inlet control:"], "x");
        if (v.size() == 2) { r.ctrl_every = (int)v[0]; r.ctrl_start = (int)v[1]; }
    }
    if (has("DYNY2:")) r.dyn_y2 = (int)num("DYNY2:");
    if (has("Dynamic inlet mode:")) r.dyn_mode = (int)num("Dynamic inlet mode:");
    if (has("Trip until:")) r.trip_until = (int)num("Trip until:");
    if (has("Trip z:")) r.trip_z = (int)num("Trip z:");
    return true;
}

// C++ ostream default formatting (6 significant digits), as WriteInputs prints
std::string g(double v)
{
    std::ostringstream o;
    o << v;
    return o.str();
}

double omega_lr_of(const Run& r)
{
    if (r.omega_lr > 0) return r.omega_lr;
    double o = r.omega;
    const int n = (int)std::lround(std::log2(1.0 / r.lr_factor));
    for (int i = 0; i < n; ++i) o = 2.0 / (1.0 + 2.0 * (2.0 / o - 1.0)); // main(), :2870-2877
    return o;
}

std::string inputs_record(const Run& r)
{
    std::ostringstream o;
    o << "Base domain size \t" << r.nx << "x" << r.ny << "x" << r.nz << "\n";
    o << "Base blocksize: \t256x1x1\n";
    o << "Obst1 location: \t(" << g(r.obst1[0]) << "," << g(r.obst1[1]) << "," << g(r.obst1[2]) << ")\n";
    o << "Obst1 radius: \t" << g(r.obst1_r) << "\n";
    o << "Obst2 location: \t(" << g(r.obst2[0]) << "," << g(r.obst2[1]) << "," << g(r.obst2[2]) << ")\n";
    o << "Obst2 radius: \t" << g(r.obst2_r) << "\n";
    o << "RE: \t" << g(r.re) << "\n";
    o << "UMAX: \t" << g(r.umax) << "\n";
    o << "omega \t: " << g(r.omega) << "\n";
    o << "DPDY \t: " << g(r.dpdy) << "\n";
    o << "TMAX: \t" << r.tmax << "\n";
    o << "STARTF: \t" << r.startf << "\n";
    o << "START_VELAV: \t" << r.start_velav << "\n";
    o << "START_VELFLUC: \t" << r.start_velfluc << "\n";
    o << "REFINEMENT: \t" << (int)r.refinement << "\n";
    o << "MODEL: \t" << r.model << "\n";
    o << "Smagorinsky LES: \t" << (int)r.les << "\n";
    o << "CS: \t" << g(r.cs) << "\n";
    o << "LR domain size \t" << r.lr[0] << "x" << r.lr[1] << "x" << r.lr[2] << "\n";
    o << "LR factor \t" << g(r.lr_factor) << "\n";
    o << "LR location \t" << g(r.lr_loc[0]) << "x" << g(r.lr_loc[1]) << "x" << g(r.lr_loc[2]) << "\n";
    o << "LR blocksize: \t256x1x1\n";
    o << "omega in LR \t: " << g(omega_lr_of(r)) << "\n";
    o << "GPUs per node \t: " << r.gpus << "\n";
    o << "# lbm-mi355x\n";
    o << "MRT rates: \t" << r.rates << "\n";
    o << "Boundary x: \t" << r.bc[0] << "\n";
    o << "Boundary y: \t" << r.bc[1] << "\n";
    o << "Boundary z: \t" << r.bc[2] << "\n";
    o << "Halfway bounce-back: \t" << (int)r.halfway << "\n";
    o << "Storage: \t" << r.storage << "\n";
    o << "Geometry: \t" << r.geometry << "\n";
    o << "Inlet perturbation: \t" << g(r.inlet_amp) << "\n";
    o << "Seed: \t" << r.seed << "\n";
    o << "Temporal interpolation: \t" << (int)r.temporal << "\n";
    o << "Force steps: \t" << (r.force_steps < 0 ? 0 : r.force_steps) << "\n";
    o << "Statistics: \t" << (int)r.velav << "\n";
    o << "DYNY1: \t" << r.dyn_y1 << "\n";
    o << "KP: \t" << g(r.kp) << "\n";
    o << "Inlet control: \t" << r.ctrl_every << "x" << r.ctrl_start << "\n";
    o << "DYNY2: \t" << r.dyn_y2 << "\n";
    o << "Dynamic inlet mode: \t" << r.dyn_mode << "\n";
    o << "Trip until: \t" << r.trip_until << "\n";
    o << "Trip z: \t" << r.trip_z << "\n";
    return o.str();
}

int code_of(const std::string& s, const char* const* names, const int* codes, int n, int fallback)
{
    for (int i = 0; i < n; ++i)
        if (s == names[i]) return codes[i];
    return fallback;
}

bool to_config(const Run& r, lbm_config& c, std::string& why)
{
    lbm_config_default(&c);
    c.nx = r.nx; c.ny = r.ny; c.nz = r.nz;
    const char* bcn[3] = {"periodic", "clamp", "wall"};
    const int bcc[3] = {LBM_BC_PERIODIC, LBM_BC_CLAMP, LBM_BC_WALL};
    for (int a = 0; a < 3; ++a) {
        c.bc[a] = code_of(r.bc[a], bcn, bcc, 3, -1);
        if (c.bc[a] < 0) { why = "unknown boundary mode '" + r.bc[a] + "'"; return false; }
    }
    if (r.model == "BGK" || r.model == "bgk") c.model = LBM_MODEL_BGK;
    else {
        const char* rn[3] = {"lrmgpu", "dhumieres", "cpu3d"};
        const int rc[3] = {LBM_MODEL_MRT_LES, LBM_MODEL_MRT_DHUMIERES, LBM_MODEL_MRT_CPU3D};
        c.model = code_of(r.rates, rn, rc, 3, -1);
        if (c.model < 0) { why = "unknown MRT rates '" + r.rates + "'"; return false; }
    }
    c.halfway = r.halfway;
    const char* sn[4] = {"f32", "f16", "bf16", "f64"};
    const int sc[4] = {LBM_STORE_F32, LBM_STORE_F16, LBM_STORE_BF16, LBM_STORE_F64};
    c.storage = code_of(r.storage, sn, sc, 4, -1);
    if (c.storage < 0) { why = "unknown storage '" + r.storage + "'"; return false; }
    const char* gn[4] = {"none", "channel", "cpu3d", "precursor"};
    const int gc[4] = {LBM_GEOM_NONE, LBM_GEOM_CHANNEL, LBM_GEOM_CPU3D, LBM_GEOM_PRECURSOR};
    c.geometry = code_of(r.geometry, gn, gc, 4, -1);
    if (c.geometry < 0) { why = "unknown geometry '" + r.geometry + "'"; return false; }
    c.omega = (float)r.omega;
    c.cs = r.les ? (float)r.cs : 0.f;
    c.dpdy = (float)r.dpdy;
    c.umax = (float)r.umax;
    c.inlet_amp = (float)r.inlet_amp;
    c.seed = (uint64_t)r.seed;
    c.obst_x = (float)r.obst1[0]; c.obst_y = (float)r.obst1[1]; c.obst_z = (float)r.obst1[2];
    c.obst_r = (float)r.obst1_r;
    c.startf = (int)r.startf;
    // WriteForces covers STARTF .. TMAX-1 (ForceTime = TMAX - STARTF)
    c.force_steps = r.force_steps >= 0 ? r.force_steps
                                       : (c.geometry != LBM_GEOM_NONE ? (int)std::max(0LL, r.tmax - r.startf) : 0);
    c.refinement = r.refinement;
    c.lr_nx = r.lr[0]; c.lr_ny = r.lr[1]; c.lr_nz = r.lr[2];
    c.lr_x0 = (float)r.lr_loc[0]; c.lr_y0 = (float)r.lr_loc[1]; c.lr_z0 = (float)r.lr_loc[2];
    c.temporal_interp = r.temporal;
    c.dyn_y1 = r.dyn_y1;
    c.kp = (float)r.kp;
    c.ctrl_every = r.ctrl_every;
    c.ctrl_start = r.ctrl_start;
    c.dyn_mode = r.dyn_mode;
    c.dyn_y2 = r.dyn_y2;
    c.obst2_x = (float)r.obst2[0]; c.obst2_y = (float)r.obst2[1]; c.obst2_z = (float)r.obst2[2];
    c.obst2_r = (float)r.obst2_r;
    c.trip_until = r.trip_until;
    c.trip_z = r.trip_z;
    return true;
}

int die(const char* what, lbm_ctx* ctx)
{
    std::fprintf(stderr, "lbm_run: %s: %s\n", what, lbm_last_error(ctx));
    if (ctx) lbm_destroy(ctx);
    return 1;
}

} // namespace

int main(int argc, char** argv)
{
    std::string inputs, casename;
    long long steps = -1;
    bool parse_only = false;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        if (a == "--case" && i + 1 < argc) casename = argv[++i];
        else if (a == "--steps" && i + 1 < argc) steps = std::atoll(argv[++i]);
        else if (a == "--parse-only") parse_only = true;
        else if (inputs.empty()) inputs = a;
        else { std::fprintf(stderr, "lbm_run: unexpected argument %s\n", a.c_str()); return 2; }
    }
    if (inputs.empty()) {
        std::fprintf(stderr, "usage: lbm_run CASE.inputs [--case NAME] [--steps N] [--parse-only]\n");
        return 2;
    }
    Run r;
    if (!parse(inputs, r)) { std::fprintf(stderr, "lbm_run: cannot read %s\n", inputs.c_str()); return 1; }
    if (steps >= 0) r.tmax = steps;
    if (parse_only) {
        std::fputs(inputs_record(r).c_str(), stdout);
        return 0;
    }
    if (casename.empty()) {
        casename = inputs;
        const size_t dot = casename.rfind(".inputs");
        if (dot != std::string::npos && dot + 7 == casename.size()) casename.resize(dot);
    }
    lbm_config c;
    std::string why;
    if (!to_config(r, c, why)) { std::fprintf(stderr, "lbm_run: %s\n", why.c_str()); return 1; }
    {
        std::ofstream o(casename + ".inputs");
        o << inputs_record(r);
    }
    lbm_ctx* ctx = nullptr;
    if (lbm_create(&c, &ctx)) return die("lbm_create", nullptr);
    if (lbm_set_geometry(ctx, nullptr)) return die("lbm_set_geometry", ctx);
    // initialize (MRTtest/DynInlet3.cu:1186-1216) for the MRT family; the
    // CPU_3D program's equilibrium fill for BGK and the CPU_3D rates
    const int init = (c.model == LBM_MODEL_MRT_LES || c.model == LBM_MODEL_MRT_DHUMIERES) ? LBM_INIT_MRT : LBM_INIT_BGK;
    if (lbm_init_equilibrium(ctx, init, nullptr, nullptr, nullptr, nullptr)) return die("lbm_init_equilibrium", ctx);
    if (r.velav && lbm_stats_enable(ctx, r.start_velav, std::max(r.start_velav, r.start_velfluc)))
        return die("lbm_stats_enable", ctx);
    lbm_sizes sz[2];
    int nlev = 1;
    if (lbm_query_sizes(ctx, 0, &sz[0])) return die("lbm_query_sizes", ctx);
    if (r.refinement && lbm_query_sizes(ctx, 1, &sz[1]) == LBM_OK) nlev = 2;
    double updates = (double)sz[0].cells + (nlev == 2 ? 2.0 * (double)sz[1].cells : 0.0);
    std::printf("lbm_run: %s  %dx%dx%d%s, %lld steps\n", casename.c_str(), r.nx, r.ny, r.nz,
                nlev == 2 ? " + fine patch" : "", r.tmax);
    std::fflush(stdout);
    // the time loop (:3119-3250) in chunks, so progress shows
    const long long chunk = std::max(1LL, r.tmax / 10);
    double gpu_ms = 0;
    const auto t0 = std::chrono::steady_clock::now();
    for (long long t = 0; t < r.tmax; t += chunk) {
        const int n = (int)std::min(chunk, r.tmax - t);
        float ms = 0;
        if (lbm_step_timed(ctx, n, &ms)) return die("lbm_step", ctx);
        gpu_ms += ms;
        std::printf("  step %lld / %lld  (%.2f GLUPS)\n", t + n, r.tmax, updates * n / (ms * 1e-3) / 1e9);
        std::fflush(stdout);
    }
    const double wall = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (lbm_write_tecplot(ctx, casename.c_str())) return die("lbm_write_tecplot", ctx);
    // WriteForces: "t, Fx/ref, Fy/ref, Fz/ref", ref = UMAX^2 ZDIM OBSTR1, times
    // LRLEVEL^2 = 4 with refinement, where the force is the fine level's
    // (WriteForces(..., REFINEMENT*LRLEVEL), MRTtest/DynInlet3.cu:2790, :3300)
    int nf = 0;
    if (lbm_get_forces(ctx, nullptr, 0, &nf)) return die("lbm_get_forces", ctx);
    if (nf > 0) {
        std::vector<double> F(3 * (size_t)nf);
        if (lbm_get_forces(ctx, F.data(), nf, &nf)) return die("lbm_get_forces", ctx);
        float ref = (float)r.umax * (float)r.umax * (float)r.nz * (float)r.obst1_r;
        if (nlev == 2) ref *= 4.f;
        std::ofstream o(casename + ".force");
        for (int i = 0; i < nf; ++i)
            o << i + r.startf << ", " << (float)F[3 * i] / ref << ", " << (float)F[3 * i + 1] / ref << ", "
              << (float)F[3 * i + 2] / ref << "\n";
    }
    // the dynamic-inlet record: "t, Av_V, dpdy" at control events with
    // t % 1000 == 0 (main(), MRTtest/DynInlet3.cu:3170), in CASENAME.vel
    {
        int nc = 0;
        if (lbm_get_inlet_control(ctx, nullptr, 0, &nc)) return die("lbm_get_inlet_control", ctx);
        std::vector<double> rows(3 * (size_t)std::max(nc, 1));
        if (nc > 0 && lbm_get_inlet_control(ctx, rows.data(), nc, &nc)) return die("lbm_get_inlet_control", ctx);
        std::ofstream o(casename + ".vel");
        for (int i = 0; i < nc; ++i)
            if ((long long)rows[3 * i] % 1000 == 0)
                o << (long long)rows[3 * i] << ", " << (float)rows[3 * i + 1] << ", " << (float)rows[3 * i + 2] << "\n";
    }
    std::printf("lbm_run: %lld steps in %.3f s GPU (%.3f s wall), %.2f GLUPS\n", r.tmax, gpu_ms * 1e-3, wall,
                r.tmax > 0 ? updates * r.tmax / (gpu_ms * 1e-3) / 1e9 : 0.0);
    lbm_destroy(ctx);
    return 0;
}
