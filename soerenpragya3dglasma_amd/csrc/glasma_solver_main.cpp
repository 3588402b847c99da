// glasma_solver: the reference's command-line simulation drivers on top of the C ABI (include/glasma.h).
//
// One program for the three drivers the reference selects at compile time (SUNcSolver.cpp:16-24):
// collision (Simulation.cpp), left-moving (LeftNucleusSimulation.cpp) and right-moving
// (RightNucleusSimulation.cpp) nucleus.  Command-line keys, defaults, the run cadence and the text
// outputs follow the reference:
//   keys      CommandlineArguments.cpp:4-116 (-o -N -Nx -Ny -Nz -a -ax -ay -az -dT -RlOverAzL -RlOverAzR -QsA
//             -MEff -Lambda -SEED -NucleusL -NucleusR), defaults LATTICE/3DGrid.cpp:20-23,
//             LATTICE/Parameters.cpp:4-48, DYNAMICS/SUNcDynamics.cpp:30;
//   run       Simulation.cpp:327-393 (charges of both nuclei from one stream, seed -SEED, scaled by Qs0A_L;
//             initial state; initial Gauss profile, unitarity, EMT map and Gauss record; Evolve(1100));
//   evolve    Simulation.cpp:212-322: every step Update; every 20 steps the bulk observables into
//             EnergyID<seed>.txt and the Gauss-law record into GausCheck<seed>.txt; every 100 steps the
//             energy-momentum profile EMTDynT<t>ID<seed>.txt; every 1000 steps the unitarity check.
//   restart   -Loaded 1 -i1 U -i2 E -i3 J -Time t (Simulation.cpp:396-421): the fields from the reference's
//             text configuration files, the configuration saved again as UDynT/EDynT/JDynT<t>ID<seed>.txt,
//             then Evolve(100).
// Project keys: -driver collision|left|right|tmd (tmd: SimulationmitPC.cpp, the collision of physical-charge
// nuclei, with -A, -centerL, -centerR, -CoMEnergy, -Qs0, -SigmaLOverAz for the Parameters it reads, evolved to
// t = 100 with configurations every 1500 steps), -QsAL (the charge scale, Parameters::Qs0A_L, which the
// reference's command line does not reach), -tmax (Evolve's end time: 1100, or 100 after -Loaded),
// -steps (stop after this many steps), -bc 0|1 (open / periodic z), -tune 0|1 (time the step's tile
// orders first), -block n (steps per glasma_update call between observable points; at most 20),
// -state_in / -state_out (an exact binary state instead of the initial conditions / at the end), -maps 1
// (the output routines no reference driver calls: CreateMapBfCollision of the initial state as
// MapBfCollisionID<seed>.txt, and with every EMT profile SaveFields / SaveTransverselyIntegratedCurrents
// as FieldsDynT<t>ID<seed>.txt / CurrentsDynT<t>ID<seed>.txt).
// Errors print the library's message and exit with status 1 (the reference exits with 0 on errors).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/glasma.h"

namespace {

struct Args {
    std::map<std::string, std::string> kv;
    bool has(const char* k) const { return kv.count(k) != 0; }
    double num(const char* k, double d) const { return has(k) ? std::atof(kv.at(k).c_str()) : d; }
    long integer(const char* k, long d) const { return has(k) ? std::atol(kv.at(k).c_str()) : d; }
    std::string str(const char* k, const char* d) const { return has(k) ? kv.at(k) : std::string(d); }
};

[[noreturn]] void die(const std::string& what, const glasma_ctx* ctx, int rc = 0) {
    std::cerr << "#ERROR " << what << " (status " << rc << "): " << glasma_last_error(ctx) << std::endl;
    std::exit(1);
}

// the reference's StringManipulation::StringCast: default stream formatting of each piece
template <class... T>
std::string cat(const T&... x) {
    std::stringstream ss;
    (ss << ... << x);
    return ss.str();
}

struct Bulk {
    double T00, TXX, TYY, TZZ, SOURCE;
};

// Observables::Bulk's T00..SOURCE from the raw sums (OBSERVABLES/BulkObservables.cpp:22-62)
Bulk bulk(glasma_ctx* ctx, const double* a) {
    double raw[9];
    if (glasma_bulk(ctx, raw)) die("glasma_bulk", ctx);
    const double acube = a[0] * a[1] * a[2];
    double t[3], src = 0.0;
    for (int m = 0; m < 3; m++) {
        const double c = (a[m] / acube) * (a[m] / acube);
        t[m] = c * raw[m] + c * raw[3 + m];
        src += a[m] / (acube * acube) * raw[6 + m];
    }
    return {0.5 * (t[0] + t[1] + t[2]), 0.5 * (t[1] + t[2] - t[0]), 0.5 * (t[0] + t[2] - t[1]),
            0.5 * (t[0] + t[1] - t[2]), src};
}

double now(const glasma_ctx* ctx) {
    double t;
    long n;
    glasma_get_time(ctx, &t, &n);
    return t;
}

long steps_done(const glasma_ctx* ctx) {
    double t;
    long n;
    glasma_get_time(ctx, &t, &n);
    return n;
}

}  // namespace

int main(int argc, char** argv) {
    Args A;
    for (int i = 1; i + 1 < argc; i += 2) {
        if (argv[i][0] != '-') {
            std::cerr << "#ERROR expected -key value pairs, got " << argv[i] << std::endl;
            return 2;
        }
        A.kv[argv[i] + 1] = argv[i + 1];
    }
    // lattice (LATTICE/3DGrid.cpp:20-23; -N sets all three, -Nx/-Ny/-Nz override)
    const long Ns = A.integer("N", -1);
    glasma_params p;
    p.nc = 3;
    const long n_def[3] = {128, 128, 2048};
    const char* nkey[3] = {"Nx", "Ny", "Nz"};
    for (int d = 0; d < 3; d++) p.n[d] = (int)A.integer(nkey[d], Ns > 0 ? Ns : n_def[d]);
    const double as = A.num("a", -1.0);
    const double a_def[3] = {1.0, 1.0, 0.25};
    const char* akey[3] = {"ax", "ay", "az"};
    for (int d = 0; d < 3; d++) p.a[d] = A.num(akey[d], as > 0 ? as : a_def[d]);
    p.dt = A.num("dT", 0.08);
    p.bc_z = (int)A.integer("bc", 0);
    // MV parameters (LATTICE/Parameters.cpp:4-28: Qs0A = 0.33, Lambda = 5 Qs0A, MEff = Qs0A at the default
    // spacings; the charges are scaled by Qs0A_L = 0.125 / sqrt(ax ay))
    const double qs0a = A.num("QsA", 0.33);
    const double qsl = A.num("QsAL", 0.125 / std::sqrt(p.a[0] * p.a[1]));
    glasma_mv_params mv;
    mv.lambda = A.num("Lambda", 5 * 0.33);
    mv.m_eff = A.num("MEff", 0.33);
    mv.sigma_l = A.num("RlOverAzL", 16.0);
    mv.sigma_r = A.num("RlOverAzR", 16.0);
    mv.nucleus_l = A.num("NucleusL", 1.0);
    mv.nucleus_r = A.num("NucleusR", 1.0);
    mv.center_l = (int)(0.85 * p.n[2] / 2 - 1);
    mv.center_r = (int)(1.15 * p.n[2] / 2 - 1);
    const long seed = A.integer("SEED", 0);
    const std::string drv = A.str("driver", "collision");
    const bool tmd = drv == "tmd";
    const int driver = drv == "left" ? 1 : (drv == "right" ? 2 : 0);
    if (driver == 0 && drv != "collision" && !tmd) {
        std::cerr << "#ERROR -driver must be collision, left, right or tmd" << std::endl;
        return 2;
    }
    // physical-charge nuclei (SimulationmitPC.cpp; LATTICE/Parameters.cpp:30-46: Qs0 = 0.5 GeV, Qs0R = 1,
    // SigmaLOverAz = 512, AtomicNumber = 197, CoMEnergy = 500 / 0.197 fm^-1; TMDbased.cpp:43,48: centres)
    glasma_tmd_params tp;
    tp.qs0a = qs0a;
    tp.m_eff = mv.m_eff;
    tp.com_energy = A.num("CoMEnergy", 500 / 0.197);
    tp.a_fm[0] = tp.a_fm[1] = qs0a * 0.197 / A.num("Qs0", 0.5);
    tp.a_fm[2] = tp.a_fm[0] * 1.0 / (qs0a * A.num("SigmaLOverAz", 512));
    tp.atomic_number = (int)A.integer("A", 197);
    tp.center_l = (int)A.integer("centerL", 785);
    tp.center_r = (int)A.integer("centerR", 1261);
    const bool loaded = A.integer("Loaded", 0) != 0;
    const std::string state_in = A.str("state_in", ""), state_out = A.str("state_out", "");
    const double tmax = A.num("tmax", (loaded || tmd) ? 100.0 : 1100.0);
    const long max_steps = A.integer("steps", -1);
    const int block = (int)std::max(1L, std::min(20L, A.integer("block", 20)));
    std::string out = A.str("o", "OUTPUT");
    std::filesystem::create_directories(out);
    if (!out.empty() && out.back() != '/') out += '/';

    std::cerr << "#GAUGE GROUP IS SU(3)" << std::endl;
    std::cerr << "## LATTICE SIZE IS " << p.n[0] << "x" << p.n[1] << "x" << p.n[2] << std::endl;
    std::cerr << "## LATTICE SPACINGS ARE " << p.a[0] << " " << p.a[1] << " " << p.a[2] << std::endl;
    std::cerr << "## DYNAMICAL TIME STEP IS " << p.dt << std::endl;
    glasma_ctx* ctx = nullptr;
    if (const int rc = glasma_create(&p, &ctx)) die("glasma_create", ctx, rc);
    std::cerr << "#STARTING ID=" << seed << " (" << drv << ")" << std::endl;

    // info file (Simulation.cpp:133-158)
    {
        std::ofstream f(cat(out, "InfoID", seed, ".txt"));
        f.precision(15);
        f << "#LATTICE DATA" << std::endl;
        f << "#Nx,Ny,Nz= " << p.n[0] << " " << p.n[1] << " " << p.n[2] << std::endl;
        f << "#ax,ay,az= " << p.a[0] << " " << p.a[1] << " " << p.a[2] << std::endl;
        f << "#dT= " << p.dt << std::endl;
        f << "    " << std::endl;
        f << "QsAperp= " << qs0a << std::endl;
        f << "Qs0A_L (charge scale)= " << qsl << std::endl;
        f << "MEff aka Infrared Cutoff= " << mv.m_eff / qs0a << " (times QsAperp)" << std::endl;
        f << "Lambda aka UV Cutoff= " << mv.lambda / qs0a << " (times QsAperp)" << std::endl;
        f << "SigmaLOverAz L= " << mv.sigma_l << std::endl;
        f << "SigmaLOverAz R= " << mv.sigma_r << std::endl;
        f << "driver= " << drv << std::endl;
    }

    const long Nz = p.n[2];
    std::vector<double> per_z(Nz), emt(10 * Nz);
    double gmax, grms, umax, urms;
    const bool fresh = !loaded && state_in.empty();
    if (loaded) {
        // Simulation::Run(f1, f2, f3, t) (Simulation.cpp:396-421)
        const std::string i1 = A.str("i1", ""), i2 = A.str("i2", ""), i3 = A.str("i3", "");
        const double T = A.num("Time", 0.0);
        auto name = [](const std::string& f) { return f.empty() ? nullptr : f.c_str(); };
        if (int rc = glasma_load_configuration(ctx, name(i1), name(i2), name(i3), T)) die("glasma_load_configuration", ctx, rc);
        const std::string u = cat(out, "UDynT", now(ctx), "ID", seed, ".txt"), e = cat(out, "EDynT", now(ctx), "ID", seed, ".txt"),
                          j = cat(out, "JDynT", now(ctx), "ID", seed, ".txt");
        if (int rc = glasma_save_configuration(ctx, u.c_str(), e.c_str(), j.c_str())) die("glasma_save_configuration", ctx, rc);
    } else if (!state_in.empty()) {
        if (int rc = glasma_load_state(ctx, state_in.c_str())) die("glasma_load_state", ctx, rc);
    } else {
        if (tmd) {
            // SimulationmitPC::Run (SimulationmitPC.cpp:326-370): noise and nucleons of both nuclei from one stream
            const size_t nn = 8 * (size_t)p.n[0] * p.n[1] * p.n[2];
            std::vector<double> nl(nn), nr(nn), pl(3 * (size_t)tp.atomic_number), pr(3 * (size_t)tp.atomic_number);
            if (glasma_tmd_draw((unsigned long)seed, p.n[0], p.n[1], p.n[2], tp.atomic_number, nl.data(), nr.data(),
                                pl.data(), pr.data()))
                die("glasma_tmd_draw (no Woods-Saxon parameters for this -A)", ctx);
            if (glasma_tmd_initial_state(ctx, 0, &tp, nl.data(), nr.data(), pl.data(), pr.data()))
                die("glasma_tmd_initial_state", ctx);
        } else {
            // initial state (Simulation.cpp:337-378)
            std::vector<double> xl(8L * p.n[0] * p.n[1]), xr(8L * p.n[0] * p.n[1]);
            if (glasma_draw_charges((unsigned long)seed, p.n[0], p.n[1], qsl, qsl, xl.data(), xr.data()))
                die("glasma_draw_charges", ctx);
            if (glasma_mv_initial_state(ctx, driver, &mv, xl.data(), xr.data())) die("glasma_mv_initial_state", ctx);
        }
    }
    if (A.integer("tune", 0)) {
        if (glasma_tune_step(ctx, 3, nullptr, nullptr, nullptr, nullptr)) die("glasma_tune_step", ctx);
    }
    auto gauss_record = [&]() {  // GaussLaw::SaveViolation("GausCheck") (GaussLawViolation.cpp:141-164)
        if (glasma_gauss(ctx, nullptr, &gmax, &grms)) die("glasma_gauss", ctx);
        std::ofstream f(cat(out, "GausCheck", seed, ".txt"), std::ios_base::app);
        f.precision(15);
        f << now(ctx) << " " << gmax << "  " << grms << std::endl;
    };
    auto emt_map = [&]() {  // EnergyMomentumTensor::CreateMap("EMTDynT<t>") (EnergyMomentumTensor.cpp:92-170)
        if (glasma_emt(ctx, emt.data())) die("glasma_emt", ctx);
        std::ofstream f(cat(out, "EMTDynT", now(ctx), "ID", seed, ".txt"));
        f.precision(15);
        for (long z = 0; z < Nz; z++) {
            f << z;
            for (int k = 0; k < 10; k++) f << "  " << emt[10 * z + k];
            f << std::endl;
        }
    };
    const bool maps = A.integer("maps", 0) != 0;
    auto profiles = [&]() {
        if (!maps) return;
        {  // IO::SaveFields (IO/SaveConfiguration.cpp:112-158)
            std::vector<double> v(4 * Nz);
            if (glasma_field_profile(ctx, v.data())) die("glasma_field_profile", ctx);
            std::ofstream f(cat(out, "FieldsDynT", now(ctx), "ID", seed, ".txt"));
            f.precision(15);
            for (long z = 0; z < Nz; z++)
                f << z << "  " << v[4 * z] << " " << v[4 * z + 1] << "  " << v[4 * z + 2] << "  " << v[4 * z + 3] << std::endl;
        }
        {  // IO::SaveTransverselyIntegratedCurrents (IO/SaveConfiguration.cpp:160-204)
            std::vector<double> v(4 * (Nz + 2));
            if (glasma_current_profile(ctx, v.data())) die("glasma_current_profile", ctx);
            std::ofstream f(cat(out, "CurrentsDynT", now(ctx), "ID", seed, ".txt"));
            f.precision(15);
            for (long r = 0; r < Nz + 2; r++)
                f << r - 1 << "  " << v[4 * r] << "  " << v[4 * r + 1] << "  " << v[4 * r + 2] << "  " << v[4 * r + 3] << std::endl;
        }
    };
    if (fresh && maps) {  // EnergyMomentumTensor::CreateMapBfCollision (EnergyMomentumTensor.cpp:10-77)
        std::vector<double> m(2L * p.n[0] * p.n[1]);
        if (glasma_transverse_energy_map(ctx, m.data())) die("glasma_transverse_energy_map", ctx);
        std::ofstream f(cat(out, "MapBfCollisionID", seed, ".txt"));
        f.precision(15);
        for (long y = 0; y < p.n[1]; y++)
            for (long x = 0; x < p.n[0]; x++) {
                const long i = 2 * (x + (long)p.n[0] * y);
                f << x << "  " << y << "  " << m[i] << "  " << m[i + 1] << std::endl;
            }
    }
    if (fresh) {
        {  // GaussLaw::CheckViolation("GVDynT<t>") per z (GaussLawViolation.cpp:16-74)
            if (glasma_gauss(ctx, per_z.data(), &gmax, &grms)) die("glasma_gauss", ctx);
            std::ofstream f(cat(out, "GVDynT", now(ctx), "ID", seed, ".txt"));
            f.precision(15);
            for (long z = 0; z < Nz; z++) f << z << "  " << per_z[z] << std::endl;
            std::cerr << "#GAUSS LAW VIOLATION AT T=" << now(ctx) << " MAX=" << gmax << " AVG=" << grms << std::endl;
        }
        if (glasma_unitarity(ctx, &umax, &urms)) die("glasma_unitarity", ctx);
        std::cerr << "#UNITARITY VIOLATION AT T=" << now(ctx) << " MAX=" << umax << " AVG=" << urms << std::endl;
        emt_map();
        profiles();
        gauss_record();
    }

    // Evolve (Simulation.cpp:212-322)
    std::ofstream energy(cat(out, "EnergyID", seed, ".txt"));
    energy.precision(15);
    {
        const Bulk b = bulk(ctx, p.a);
        energy << now(ctx) << " " << b.T00 << " " << b.TXX << " " << b.TYY << " " << b.TZZ << " " << b.SOURCE << std::endl;
    }
    std::cerr << "#STARTING TIME EVOLUTION" << std::endl;
    while (now(ctx) < tmax && (max_steps < 0 || steps_done(ctx) < max_steps)) {
        // advance to the next observable point (a multiple of `block` steps), or by one step near the end
        long n = block - steps_done(ctx) % block;
        if (max_steps >= 0) n = std::min(n, max_steps - steps_done(ctx));
        if (now(ctx) + (n - 1) * p.dt >= tmax) n = 1;
        if (glasma_update(ctx, (int)n)) die("glasma_update", ctx);
        const long s = steps_done(ctx);
        if (s % 20 == 0) {
            const Bulk b = bulk(ctx, p.a);
            std::cerr << s << "  " << now(ctx) << " " << b.T00 << " " << b.SOURCE << std::endl;
            energy << s << "  " << now(ctx) << "  " << b.T00 << "  " << b.SOURCE << std::endl;
            gauss_record();
        }
        if (s % 100 == 0) {
            std::cerr << "#SAVING OBSERVABLES AT T=" << s << "  " << now(ctx) << std::endl;
            emt_map();
            profiles();
        }
        if (tmd && s % 1500 == 0) {  // SimulationmitPC.cpp Evolve: configurations every 1500 steps
            const std::string u = cat(out, "UDynT", now(ctx), "ID", seed, ".txt"),
                              e = cat(out, "EDynT", now(ctx), "ID", seed, ".txt"),
                              j = cat(out, "JDynT", now(ctx), "ID", seed, ".txt");
            if (int rc = glasma_save_configuration(ctx, u.c_str(), e.c_str(), j.c_str()))
                die("glasma_save_configuration", ctx, rc);
        }
        if (s % 1000 == 0) {
            if (tmd) {  // and the Gauss-law profile with the unitarity check every 1000 steps
                if (glasma_gauss(ctx, per_z.data(), &gmax, &grms)) die("glasma_gauss", ctx);
                std::ofstream f(cat(out, "GVDynT", now(ctx), "ID", seed, ".txt"));
                f.precision(15);
                for (long z = 0; z < Nz; z++) f << z << "  " << per_z[z] << std::endl;
            }
            if (glasma_unitarity(ctx, &umax, &urms)) die("glasma_unitarity", ctx);
            std::cerr << "#UNITARITY VIOLATION AT T=" << now(ctx) << " MAX=" << umax << " AVG=" << urms << std::endl;
        }
    }
    energy.close();
    if (!state_out.empty()) {
        if (int rc = glasma_save_state(ctx, state_out.c_str())) die("glasma_save_state", ctx, rc);
    }
    std::cerr << "#COMPLETED ID=" << seed << " after " << steps_done(ctx) << " steps, t=" << now(ctx) << std::endl;
    glasma_destroy(ctx);
    return 0;
}
