// conan_cli.cpp -- the `conan` command-line tool of this package: the ic, run,
// cic and fof commands of conan (src/ic/main.cc, src/run/main.cc,
// src/cic/main.cc, src/fof/fof.cc) with the same options, defaults, parameter
// headers and output files, computing on the GPU through the C ABI of
// include/conanpm.h.
//
// Options added to the reference's (additive only):
//   ic  --noise FILE     read the white noise from record "noise" of FILE
//   run --precision P    fp64 (reference arithmetic, default) or fp32
//   run --wrap-gather    fold positions into the box before the gather
//   run --midpoint       drift and kick factors at the middle of each step
//   run --resume FILE    continue from snapshot FILE (<id>.pos.<k>.conan)
#include <cstring>

#include "conan_format.hpp"
#include "conanpm.h"

#include <cmath>
#include <cstdlib>
#include <fstream>
#include <iomanip>
#include <sstream>
#include <functional>
#include <iostream>
#include <map>
#include <memory>
#include <random>

using namespace conan;

namespace {

void pm_ok(int rc, pm_handle* h = nullptr) {
    if (rc != 0) throw Failure{std::string("GPU: ") + pm_last_error(h)};
}

std::string timed_seed() { return fmt(time(nullptr)); }

std::string vec_dtype(int dim) { return dim == 3 ? "3f8" : dim == 2 ? "2f8" : ""; }

void help_and_exit(const Argv& C) {
    std::cout << "conan (MI355X particle-mesh build).\n\n";
    C.print(std::cout);
    exit(0);
}

// ------------------------------------------------------------------ ic
void cmd_ic(int argc, char** argv) {
    using O = Option;
    Argv C = read_arguments(
        argc, argv,
        {{0, "h", "help", "false", "print help on the use of this program."},
         {O::VALUED | O::CHECK, "i", "id", date_string(), "identifier for filenames."},
         {O::VALUED | O::CHECK, "d", "dim", "3",
          "number of dimensions in which to run the code. Usually this should by 2 or 3."},
         {O::VALUED | O::CHECK, "b", "mbits", "8",
          "size of the mass-box, or binary power of the number of particles per dimension. "
          "A value of 8 (default) will give a boxsize of 256."},
         {O::VALUED | O::CHECK, "L", "size", "100", "physical size of the simulation in units of Mpc/h."},
         {O::VALUED | O::CHECK, "", "seed", timed_seed(),
          "random seed used to generate the initial conditions. By default the number of seconds since "
          "Epoch is used. "},
         {O::VALUED | O::CHECK, "", "Omega0", "0.32", "matter density."},
         {O::VALUED | O::CHECK, "", "H0", "69.0", "Hubble parameter."},
         {O::VALUED | O::CHECK, "n", "ns", "0.96", "the slope of the power spectrum."},
         {0, "", "smooth", "false", "whether to smooth the initial conditions with a Gaussian filter or not."},
         {0, "", "scale-free", "false", "this will enable a scale-free power-spectrum with power law index <ns>."},
         {O::VALUED | O::CHECK, "", "scale", "8.0", "scale at which to smooth in units of Mpc/h."},
         {O::VALUED | O::CHECK, "", "sigma", "0.83", "normalisation at given scale."},
         {0, "p", "potential", "false", "include the potential in the result."},
         {0, "z", "displacement", "false", "include the Zel'dovich displacement in the result."},
         {O::VALUED | O::CHECK, "", "noise", "",
          "read the white noise from record 'noise' of this .conan file instead of drawing it from --seed."},
         {0, "", "nyquist-cutoff", "false",
          "leave out the modes beyond the Nyquist wave number (|k| > pi N / L) in the initial conditions."}});
    if (C.get<bool>("help")) help_and_exit(C);

    Header H;
    H << C;
    H["N"] = fmt(1 << H.get<unsigned>("mbits"));
    History I;
    I << C;

    const int dim = H.get<int>("dim");
    const int mbits = H.get<int>("mbits");
    const double L = H.get<double>("size");
    if (dim != 2 && dim != 3) throw Failure{"only 2 and 3 dimensions supported."};
    const size_t N = size_t(1) << mbits;
    size_t size = 1;
    for (int k = 0; k < dim; ++k) size *= N;

    // white noise (Gaussian_white_noise, src/base/misc.hh:179-188; ic.cc:55-57)
    std::vector<double> noise;
    if (!C["noise"].empty()) {
        std::ifstream nf(C["noise"], std::ios::binary);
        if (!nf) throw Failure{"Could not read from file"};
        Header NH(nf);
        History NI(nf);
        noise = load_array<double>(nf, "noise");
        if (noise.size() != size) throw Failure{"noise record has the wrong size"};
    } else {
        std::mt19937 rng(H.get<unsigned>("seed"));
        std::normal_distribution<double> normal;
        noise.resize(size);
        for (auto& v : noise) v = normal(rng);
    }

    std::vector<double> D(size);
    if (H.get<bool>("scale-free")) {
        pm_ic_params p{dim, mbits, L, H.get<double>("ns"), H.get<bool>("smooth") ? 1 : 0, H.get<double>("scale"),
                       H.get<double>("sigma"), H.get<bool>("nyquist-cutoff") ? 1 : 0};
        pm_ok(pm_ic_scale_free(&p, noise.data(), D.data(), nullptr));
    } else {
        pm_ok(pm_ic_cdm(dim, mbits, L, H.get<double>("ns"), H.get<double>("Omega0"), H.get<double>("H0"),
                        H.get<bool>("nyquist-cutoff") ? 1 : 0, noise.data(), D.data()));
    }

    std::ofstream fo(C["id"] + ".density.init.conan", std::ios::binary);
    H.to_file(fo);
    I.to_file(fo);
    save_record(fo, "density", "float64", D.data(), D.size() * sizeof(double));
    if (C.get<bool>("potential")) {
        pm_ok(pm_ic_potential(dim, mbits, L, D.data()));
        save_record(fo, "potential", "float64", D.data(), D.size() * sizeof(double));
    }
    if (C.get<bool>("displacement")) {
        // computed from D: the potential when -p was given, else the density (ic/main.cc:81-90)
        std::vector<double> psi(size * dim);
        pm_ok(pm_ic_displacement(dim, mbits, L, D.data(), psi.data()));
        save_record(fo, "displacement", vec_dtype(dim), psi.data(), psi.size() * sizeof(double));
    }
}

// ------------------------------------------------------------------ run
enum Format { FMT_ASCII, FMT_IFRIT, FMT_CONAN };

struct Snapshotter {
    pm_handle* h;
    int dim;
    unsigned N;
    double L;
    std::string id;
    Format format;

    // Gravity::save_snapshot / save_conan / save_ifrit (src/run/nbody.hh:135-218)
    void save(unsigned i, double a) {
        if ((i + 1) % 10 != 0) return;
        std::cout << "|";
        std::cout.flush();
        size_t np = 1;
        for (int k = 0; k < dim; ++k) np *= N;
        std::vector<double> X(np * dim), P(np * dim);
        pm_ok(pm_get_state(h, X.data(), P.data()), h);
        double apos = 0, amom = 0;
        pm_ok(pm_get_times(h, &apos, &amom), h);
        const unsigned k = (i + 1) / 10;
        std::string fn = id + ".pos." + fmt(k) + (format == FMT_ASCII ? ".txt" : format == FMT_IFRIT ? ".bin" : ".conan");
        std::ofstream fo(fn, std::ios::binary);
        if (format == FMT_ASCII) {
            for (size_t p = 0; p < np; ++p) {
                for (int c = 0; c < dim; ++c) fo << X[p * dim + c] << " ";
                fo << " ";
                for (int c = 0; c < dim; ++c) fo << P[p * dim + c] / a << " ";
                fo << "\n";
            }
        } else if (format == FMT_IFRIT) {
            if (dim != 3) throw Failure{"IFRIT files only support 3D"};
            uint32_t n32 = (uint32_t)np;
            write_block_32(fo, &n32, sizeof n32);
            float box[6] = {0, 0, 0, (float)L, (float)L, (float)L};
            write_block_32(fo, box, sizeof box);
            std::vector<float> f(np);
            for (const std::vector<double>* A : {&X, &P})
                for (int c = 0; c < 3; ++c) {
                    for (size_t p = 0; p < np; ++p) f[p] = (float)(*A)[p * 3 + c];
                    write_block_32(fo, f.data(), (uint32_t)(np * sizeof(float)));
                }
        } else {
            Header S;
            S["id"] = id;
            S["N"] = fmt(N);
            S["L"] = fmt(L);
            S["dim"] = fmt(dim);
            S["a-mom"] = fmt(amom);
            S["a-pos"] = fmt(apos);
            double pq = 0;
            pm_ok(pm_get_momentum_quantum(h, &pq), h);
            if (pq > 0) {  // FP32 form: the momentum grid, exactly (a power of two)
                int e = 0;
                (void)std::frexp(pq, &e);
                S["momentum-quantum-log2"] = fmt(e - 1);
            }
            History Hs;
            Hs.update("ran nbody");
            S.to_file(fo);
            Hs.to_file(fo);
            save_record(fo, "pos", vec_dtype(dim), X.data(), X.size() * sizeof(double));
            save_record(fo, "mom", vec_dtype(dim), P.data(), P.size() * sizeof(double));
        }
    }
};

void cmd_run(int argc, char** argv) {
    using O = Option;
    Argv C = read_arguments(
        argc, argv,
        {{0, "h", "help", "false", "print help on the use of this program."},
         {O::VALUED | O::CHECK, "i", "id", date_string(), "identifier for filenames."},
         {O::VALUED | O::CHECK, "H0", "H0", "68.5",
          "Hubble parameter, default is point of overlap between WMAP9 and Planck."},
         {O::VALUED | O::CHECK, "m", "Omega_m", "1.0", "Omega-matter. The default universe is EdS."},
         {O::VALUED | O::CHECK, "L", "Omega_L", "0.0", "Omega-Lambda. The default universe is EdS."},
         {0, "c", "concordance", "false",
          "Override cosmological parameters to have concordance (Planck) model: H0 = 68.5, Omega_L = 0.69, "
          "Omega_m = 0.31"},
         {O::VALUED | O::CHECK, "f", "f", "2",
          "fractional force resolution. The default makes the force grid twice the size of the mass grid."},
         {O::VALUED | O::CHECK, "a0", "a0", "0.02", "start time."},
         {O::VALUED | O::CHECK, "da", "da", "0.02", "size of time step."},
         {O::VALUED | O::CHECK, "n", "n", "50", "number of iterations."},
         {0, "", "flip-flops", "false", "count number of flip-flops for each particle."},
         {0, "", "ifrit", "false", "give output in Ifrit readable format."},
         {0, "", "ascii", "false", "give output in ASCII (not recommended for 3D)."},
         {O::VALUED | O::CHECK, "", "precision", "fp64",
          "arithmetic of the PM step: fp64 (the reference's arithmetic) or fp32 (32-bit fixed-point "
          "positions, FP32 mesh, FFT and gather, FP64 momenta)."},
         {0, "", "wrap-gather", "false",
          "fold positions into the box before the force interpolation (the reference truncates the "
          "unwrapped coordinate)."},
         {0, "", "serial-deposit", "false",
          "fp64: sum each cell of the deposit in particle order like the reference (bit-identical density) "
          "instead of in exact fixed point."},
         {0, "", "midpoint", "false",
          "evaluate the drift and kick factors at the middle of each step, a second-order leapfrog (the "
          "reference evaluates them at the start of the step)."},
         {O::VALUED | O::CHECK, "", "resume", "",
          "continue from snapshot FILE (<id>.pos.<k>.conan of a run with the same parameters): its "
          "positions, momenta and momentum grid are loaded and the run goes on after step 10 k."}});
    if (C.get<bool>("help")) help_and_exit(C);

    std::ifstream fi(C["id"] + ".density.init.conan", std::ios::binary);
    Header H(fi);
    H << C;
    History I(fi);
    I << C;
    std::vector<double> phi = load_array<double>(fi, "potential");
    fi.close();

    const unsigned f = H.get<unsigned>("f"), N = H.get<unsigned>("N");
    const double L = H.get<double>("size");
    double H0 = H.get<double>("H0"), Om = H.get<double>("Omega_m"), OL = H.get<double>("Omega_L");
    if (C.get<bool>("concordance")) {
        H0 = 68.5;
        Om = 0.31;
        OL = 0.69;
    }
    const double a0 = H.get<double>("a0"), da = H.get<double>("da");
    const unsigned n = (unsigned)H.get<double>("n");
    Format format = FMT_CONAN;
    if (H.get<bool>("ifrit")) format = FMT_IFRIT;
    if (H.get<bool>("ascii")) format = FMT_ASCII;
    const int dim = H.get<int>("dim");
    if (dim != 2 && dim != 3) throw Failure{"only 2-D and 3-D runs are supported"};
    const std::string prec = H["precision"];
    if (prec != "fp64" && prec != "fp32") throw Failure{"--precision must be fp64 or fp32"};
    size_t np = 1;
    for (int k = 0; k < dim; ++k) np *= N;
    if (phi.size() != np) throw Failure{"potential record does not match N^dim"};

    pm_config cfg{dim, (int)N, (int)(f * N), L, H0, Om, OL, prec == "fp64" ? PM_FP64 : PM_FP32,
                  (H.get<bool>("wrap-gather") ? PM_FLAG_WRAP_GATHER : 0u) |
                  (H.get<bool>("serial-deposit") ? PM_FLAG_SERIAL_DEPOSIT : 0u) |
                      (H.get<bool>("midpoint") ? PM_FLAG_MIDPOINT : 0u)};
    pm_handle* h = nullptr;
    pm_ok(pm_create(&cfg, &h));
    std::unique_ptr<pm_handle, void (*)(pm_handle*)> guard(h, pm_destroy);
    Snapshotter snap{h, dim, N, L, H["id"], format};

    // LeapFrog::run (src/run/leapfrog.hh:33-46); kicks between snapshots run
    // as one fused sequence.
    double a = a0;
    unsigned i = 0;
    const std::string resume = H["resume"];
    if (resume.empty()) {
        pm_ok(pm_init_zeldovich(h, phi.data(), a0, a0 + da / 2), h);
    } else {
        // snapshot k holds the state after step 10 k (save_snapshot, src/run/nbody.hh:186-208)
        const size_t at = resume.rfind(".pos."), dot = resume.rfind(".conan");
        if (at == std::string::npos || dot == std::string::npos || dot <= at + 5)
            throw Failure{"--resume needs a snapshot named <id>.pos.<k>.conan"};
        const unsigned k = parse<unsigned>(resume.substr(at + 5, dot - at - 5));
        std::ifstream rf(resume, std::ios::binary);
        if (!rf) throw Failure{"Could not read from file"};
        Header RH(rf);
        History RI(rf);
        if (RH.get<unsigned>("N") != N || RH.get<int>("dim") != dim || RH.get<double>("L") != parse<double>(fmt(L)))
            throw Failure{"the snapshot does not match this run's N, dim and L"};
        std::vector<double> X = load_array<double>(rf, "pos"), P = load_array<double>(rf, "mom");
        if (X.size() != np * dim || P.size() != np * dim) throw Failure{"snapshot records have the wrong size"};
        if (k == 0 || 10 * k > n) throw Failure{"the snapshot index is outside this run"};
        double alast = a0;  // a of the last step in the snapshot, accumulated as LeapFrog::run does
        for (i = 0; i < 10 * k; ++i) {
            alast = a;
            a += da;
        }
        if (std::fabs(RH.get<double>("a-pos") - alast) > 1e-5 * std::fabs(alast))
            throw Failure{"the snapshot's a-pos does not belong to this run's a0 and da"};
        if (RH.has("momentum-quantum-log2")) {  // before the state: pm_set_state keeps it
            if (prec != "fp32") throw Failure{"the snapshot was written by the fp32 form"};
            pm_ok(pm_set_momentum_quantum(h, std::ldexp(1.0, RH.get<int>("momentum-quantum-log2"))), h);
        } else if (prec == "fp32") {
            throw Failure{"the snapshot has no momentum grid (not written by the fp32 form)"};
        }
        pm_ok(pm_set_state(h, X.data(), P.data()), h);
    }
    while (i < n) {
        unsigned chunk = std::min<unsigned>(n - i, 10 - (i % 10));
        pm_ok(pm_steps(h, a, da, (int)chunk, &a), h);
        std::cout << std::string(chunk, '.');
        std::cout.flush();
        i += chunk;
        snap.save(i - 1, a);
    }
}

// ------------------------------------------------------------------ cic
void cmd_cic(int argc, char** argv) {
    using O = Option;
    Argv C = read_arguments(
        argc, argv,
        {{0, "h", "help", "false", "print help on the use of this program."},
         {O::VALUED | O::CHECK, "i", "id", date_string(), "identifier for filenames."},
         {O::VALUED | O::CHECK, "f", "f", "2",
          "fractional Eulerian resolution. The default makes the Eulerian grid twice the size of the mass grid."},
         {O::VALUED | O::CHECK, "m", "m", "5", "snap-shot"},
         {O::VALUED | O::CHECK, "s", "sub-sample", "1", "subsample the grid and interpolate the position."}});
    if (C.get<bool>("help")) help_and_exit(C);

    const unsigned m = C.get<unsigned>("m");
    std::ifstream fi(C["id"] + ".pos." + fmt(m) + ".conan", std::ios::binary);
    Header H(fi);
    H << C;
    History I(fi);
    I << C;
    const unsigned f = H.get<unsigned>("f"), N = H.get<unsigned>("N");
    const double L = H.get<double>("L");
    const unsigned s = H.get<unsigned>("sub-sample");
    if (!H.has("dim")) H["dim"] = "3";

    std::ofstream fo(C["id"] + ".rho." + fmt(m) + ".conan", std::ios::binary);
    H["N"] = fmt(N * f);
    H.to_file(fo);
    I.to_file(fo);

    const int dim = H.get<int>("dim");
    if (dim != 2 && dim != 3) throw Failure{"only 2-D and 3-D meshes are supported"};
    std::vector<double> X = load_array<double>(fi, "pos");
    size_t nf = 1;
    for (int k = 0; k < dim; ++k) nf *= (size_t)N * f;
    std::vector<double> rho(nf);
    pm_ok(pm_cic(dim, (int)N, (int)(N * f), L, (int)s, X.data(), rho.data()));
    save_record(fo, "density", "float64", rho.data(), rho.size() * sizeof(double));
}

// ------------------------------------------------------------------ fof
// timed_filename (src/base/date.cc:34-50): id.stage.NNNNN.ext, NNNNN =
// round(b * 10000) of the single-precision time, zero-padded to 5 digits.
std::string timed_filename(const std::string& id, const std::string& stage, float b, const std::string& ext) {
    std::ostringstream s;
    if (b < 0.0)
        s << id << "." << stage << ".init" << ext;
    else
        s << id << "." << stage << "." << std::setfill('0') << std::setw(5) << static_cast<int>(std::round(b * 10000))
          << ext;
    return s.str();
}

// System::modulus for doubles (src/base/misc.hh:144-148)
double modulus(double a, double b) {
    if (a < 0) return a - static_cast<int>(a / b - 1) * b;
    return a - static_cast<int>(a / b) * b;
}

// random_colour (src/fof/fof.cc:150-154)
int random_colour(size_t seed) {
    srand((unsigned)seed);
    return rand() % 1022;
}

// cmd_fof (src/fof/fof.cc:244-330) with fof_run's output (fof.cc:216-241):
// the groups come from pm_fof on the GPU.
void cmd_fof(int argc, char** argv) {
    using O = Option;
    Argv C = read_arguments(
        argc, argv,
        {{0, "h", "help", "false", "print help on the use of this program."},
         {O::VALUED | O::CHECK, "i", "id", date_string(), "identifier for filenames."},
         {O::VALUED | O::CHECK, "t", "time", "1.0", "time of snapshot."},
         {O::VALUED | O::CHECK, "l", "linking-length", "0.1", "linking length in units of Mpc h^-1."},
         {O::VALUED | O::CHECK, "f", "filename", "", "if given, supersedes internal naming system."},
         {0, "", "ascii", "true", "give output in ASCII."}});
    if (C.get<bool>("help")) help_and_exit(C);

    const double t = C.get<double>("time");
    const std::string fi_fn = C["filename"] == "" ? timed_filename(C["id"], "pos", (float)t, ".conan") : C["filename"];
    std::ifstream fi(fi_fn, std::ios::binary);
    if (!fi) throw Failure{"could not open " + fi_fn};
    Header H(fi);
    H << C;
    History I(fi);
    I << C;
    const unsigned R = H.has("dim") ? H.get<unsigned>("dim") : 3;
    (void)H.get<unsigned>("N");
    const double L = H.get<double>("L"), ll = H.get<double>("linking-length");
    if (R != 2 && R != 3) throw Failure{"only 2-D and 3-D snapshots are supported"};

    std::ofstream fo(timed_filename(C["id"], "fof", (float)t, ".txt"));
    H.to_txt_file(fo);
    I.to_txt_file(fo);

    std::vector<double> X = load_array<double>(fi, "pos");
    const size_t n = X.size() / R;
    std::vector<long long> group(n);
    long long ng = 0;
    std::cerr << "[friends-of-friends on the GPU] ";
    if (pm_fof((int)R, (long long)n, L, ll, X.data(), group.data(), &ng) != 0)
        throw Failure{std::string("GPU: ") + pm_fof_last_error(nullptr)};
    std::cerr << "[found " << ng << " groups] [post-processing] ";

    // particles of each group in index order (fof.cc:218-230)
    std::vector<size_t> start(ng + 1, 0), order(n);
    for (size_t i = 0; i < n; ++i) ++start[group[i] + 1];
    for (long long g = 0; g < ng; ++g) start[g + 1] += start[g];
    {
        std::vector<size_t> fill(start.begin(), start.end() - 1);
        for (size_t i = 0; i < n; ++i) order[fill[group[i]]++] = i;
    }
    auto put = [&](const double* p) {
        for (unsigned k = 0; k < R; ++k) fo << p[k] << " ";
    };
    for (long long g = 0; g < ng; ++g) {
        const size_t mass = start[g + 1] - start[g];
        if (mass < 10) continue;
        const double* p = &X[order[start[g]] * R];
        double w[3];
        for (unsigned k = 0; k < R; ++k) w[k] = modulus(p[k], L);
        put(w);
        fo << " " << mass << " " << g << std::endl;
    }
    fo << "\n\n";
    for (long long g = 0; g < ng; ++g) {
        const int colour = random_colour((size_t)g);
        for (size_t q = start[g]; q < start[g + 1]; ++q) {
            put(&X[order[q] * R]);
            fo << " " << colour << std::endl;
        }
    }
    std::cerr << "[done]\n";
}

const std::map<std::string, std::function<void(int, char**)>>& commands() {
    static const std::map<std::string, std::function<void(int, char**)>> m = {
        {"cic", cmd_cic}, {"fof", cmd_fof}, {"ic", cmd_ic}, {"run", cmd_run}};
    return m;
}

[[noreturn]] void usage(const std::string& cmd, int status) {
    std::string names;
    for (const auto& kv : commands()) names += (names.empty() ? "" : ", ") + kv.first;
    std::cout << "conan -- particle-mesh N-body toolchain, MI355X build.\n\n"
              << "usage: " << cmd << " [command] [args]\n"
              << "where [command] is one of: " << names << "\n"
              << "to get help on a command: " << cmd << " [command] --help\n";
    exit(status);
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 2) usage(argv[0], 0);
    const std::string command(argv[1]);
    try {
        auto it = commands().find(command);
        if (it == commands().end()) throw Failure{"command [" + command + "] not understood."};
        it->second(argc - 1, argv + 1);
    } catch (const Failure& e) {
        std::cerr << "[error] " << e.msg << std::endl;
        usage(argv[0], 255);
    }
    return 0;
}
