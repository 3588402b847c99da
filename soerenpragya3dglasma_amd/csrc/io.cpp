// Configuration files: the reference's text format (IO::SaveConfiguration / SaveConfigurationJ,
// IO/SaveConfiguration.cpp:6-93, and IO::LoadConfiguration, IO/LoadConfiguration.cpp:3-202) and an exact
// binary state file for bit-identical restarts (no counterpart in the reference).
//
// Text format, one line per record, precision 15 (Definitions.cpp:33-34):
//   U file  "x y z mu" + the 9 complex entries of U_mu(x) in storage order (i + 3j) as "re im ";
//   E file  "x y z mu e^0 ... e^7";
//   J file  "x y z a J+stat^a J+^a J-stat^a J-^a ";
// records in z, y, x (, mu) order; z runs over the ghost layers -1 .. Nz (U and E of periodic lattices:
// 0 .. Nz-1, as the reference saves them).  Loading checks the record order and count as the reference
// does, accepts either z range for U and E, and reports errors instead of exiting.  Two reference bugs
// are fixed: LoadCurrents' loop `for(aIn==0;aIn<3;aIn++)` wrote the line's values into colours a..2 (an
// SU(2) remnant); each line now sets its own colour a only.  Dynamics::LoadedTime computed tSteps from the
// old time before assigning the new one; the step count is now round(t / dt).
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <sstream>
#include <string>
#include <vector>

#include "internal.h"

using namespace glasma;

namespace {

struct Host {
    std::vector<double> U, UOld, E, Jp, Jm, JpS, JmS;
};

long zlayers(const glasma_ctx* c) { return (long)c->g.nz + 2; }
long plane(const glasma_ctx* c) { return (long)c->g.nx * c->g.ny; }

int alloc(glasma_ctx* c, Host& h) {
    const long n = plane(c) * zlayers(c);
    try {
        h.U.assign(18 * 3 * n, 0.0);
        h.UOld.assign(18 * 3 * n, 0.0);
        h.E.assign(8 * 3 * n, 0.0);
        for (auto* v : {&h.Jp, &h.Jm, &h.JpS, &h.JmS}) v->assign(8 * n, 0.0);
    } catch (const std::bad_alloc&) {
        return set_error(c, GLASMA_ERR_NOMEM, "host buffers for the configuration");
    }
    return 0;
}

int io_fail(glasma_ctx* c, const std::string& m) { return set_error(c, GLASMA_ERR_IO, m); }

// site index with ghosts (LATTICE/GaugeLinks.cpp:27-36)
long site(const glasma_ctx* c, long x, long y, long z) { return x + c->g.nx * (y + (long)c->g.ny * (z + 1)); }

void z_range(const glasma_ctx* c, bool with_ghosts, long& z0, long& z1) {
    z0 = with_ghosts ? -1 : 0;
    z1 = with_ghosts ? c->g.nz : c->g.nz - 1;
}

int save_text(glasma_ctx* c, const char* uf, const char* ef, const char* jf) {
    Host h;
    int rc = alloc(c, h);
    if (rc) return rc;
    if ((rc = glasma_get_fields(c, h.U.data(), h.E.data(), h.Jp.data(), h.Jm.data(), h.JpS.data(), h.JmS.data())))
        return rc;
    long z0, z1;
    z_range(c, c->g.bc == 0, z0, z1);  // open: ghosts included (SaveConfiguration.cpp:52-62)
    if (uf || ef) {
        std::ofstream U, E;
        if (uf) U.open(uf);
        if (ef) E.open(ef);
        if ((uf && !U) || (ef && !E)) return io_fail(c, "cannot open configuration file for writing");
        U.precision(15);
        E.precision(15);
        for (long z = z0; z <= z1; z++)
            for (long y = 0; y < c->g.ny; y++)
                for (long x = 0; x < c->g.nx; x++)
                    for (int mu = 0; mu < 3; mu++) {
                        const long s = site(c, x, y, z);
                        if (uf) {
                            U << x << " " << y << " " << z << " " << mu << " ";
                            const double* m = &h.U[18 * (mu + 3 * s)];
                            for (int k = 0; k < 9; k++) U << m[2 * k] << " " << m[2 * k + 1] << " ";
                            U << "\n";
                        }
                        if (ef) {
                            E << x << " " << y << " " << z << " " << mu;
                            for (int a = 0; a < 8; a++) E << " " << h.E[a + 8 * (mu + 3 * s)];
                            E << "\n";
                        }
                    }
        if ((uf && !U) || (ef && !E)) return io_fail(c, "write error in configuration file");
    }
    if (jf) {  // SaveConfigurationJ: always z = -1 .. Nz (SaveConfiguration.cpp:73-93)
        std::ofstream J(jf);
        if (!J) return io_fail(c, std::string("cannot open ") + jf + " for writing");
        J.precision(15);
        for (long z = -1; z <= c->g.nz; z++)
            for (long y = 0; y < c->g.ny; y++)
                for (long x = 0; x < c->g.nx; x++)
                    for (int a = 0; a < 8; a++) {
                        const long i = a + 8 * site(c, x, y, z);
                        J << x << " " << y << " " << z << " " << a << " " << h.JpS[i] << " " << h.Jp[i] << " "
                          << h.JmS[i] << " " << h.Jm[i] << " \n";
                    }
        if (!J) return io_fail(c, std::string("write error in ") + jf);
    }
    return 0;
}

// records of U (R = 18 doubles after "x y z mu") or E (R = 8): in z, y, x, mu order over -1..Nz or 0..Nz-1
int load_links(glasma_ctx* c, const char* fname, int R, std::vector<double>& out) {
    std::ifstream in(fname);
    if (!in) return io_fail(c, std::string("cannot open ") + fname);
    std::string line;
    long count = 0, zstart = 0;
    bool first = true;
    std::vector<double> v(R);
    while (std::getline(in, line)) {
        if (line.find_first_not_of(" \t\r") == std::string::npos) continue;
        std::istringstream ss(line);
        long x, y, z, mu;
        if (!(ss >> x >> y >> z >> mu)) return io_fail(c, std::string("malformed record in ") + fname);
        if (first) {
            zstart = z;
            first = false;
            if (zstart != -1 && zstart != 0) return io_fail(c, std::string(fname) + ": first layer must be z = -1 or 0");
        }
        // the record the position count expects (LoadConfiguration.cpp:49-57)
        const long ex_mu = count % 3, ex_x = (count / 3) % c->g.nx, ex_y = (count / 3 / c->g.nx) % c->g.ny,
                   ex_z = count / 3 / plane(c) + zstart;
        if (x != ex_x || y != ex_y || z != ex_z || mu != ex_mu) {
            std::ostringstream m;
            m << fname << ": positions do not match at record " << count << " (read " << x << " " << y << " " << z
              << " " << mu << ")";
            return io_fail(c, m.str());
        }
        if (z > c->g.nz) return io_fail(c, std::string(fname) + ": more layers than the lattice");
        for (int k = 0; k < R; k++)
            if (!(ss >> v[k])) return io_fail(c, std::string("could not read the values of a record in ") + fname);
        const long s = site(c, x, y, z);
        for (int k = 0; k < R; k++) out[k + R * (mu + 3 * s)] = v[k];
        count++;
    }
    const long want = 3 * plane(c) * (zstart == -1 ? zlayers(c) : c->g.nz);
    if (count != want) {
        std::ostringstream m;
        m << fname << ": " << count << " records, expected " << want;
        return io_fail(c, m.str());
    }
    return 0;
}

int load_currents(glasma_ctx* c, const char* fname, Host& h) {
    std::ifstream in(fname);
    if (!in) return io_fail(c, std::string("cannot open ") + fname);
    std::string line;
    long count = 0;
    while (std::getline(in, line)) {
        if (line.find_first_not_of(" \t\r") == std::string::npos) continue;
        std::istringstream ss(line);
        long x, y, z, a;
        double jps, jp, jms, jm;
        if (!(ss >> x >> y >> z >> a >> jps >> jp >> jms >> jm)) return io_fail(c, std::string("malformed record in ") + fname);
        if (x < 0 || x >= c->g.nx || y < 0 || y >= c->g.ny || z < -1 || z > c->g.nz || a < 0 || a >= 8)
            return io_fail(c, std::string(fname) + ": position outside the lattice");
        const long i = a + 8 * site(c, x, y, z);
        h.JpS[i] = jps;
        h.Jp[i] = jp;
        h.JmS[i] = jms;
        h.Jm[i] = jm;
        count++;
    }
    const long want = 8 * plane(c) * zlayers(c);
    if (count != want && count != 8 * plane(c) * c->g.nz) {
        std::ostringstream m;
        m << fname << ": " << count << " records, expected " << want;
        return io_fail(c, m.str());
    }
    return 0;
}

constexpr char kMagic[8] = {'G', 'L', 'A', 'S', 'M', 'A', 'S', '1'};

}  // namespace

extern "C" {

int glasma_save_configuration(glasma_ctx* ctx, const char* ufile, const char* efile, const char* jfile) {
    if (!ctx) return GLASMA_ERR_ARG;
    return save_text(ctx, ufile, efile, jfile);
}

int glasma_load_configuration(glasma_ctx* ctx, const char* ufile, const char* efile, const char* jfile, double t) {
    if (!ctx) return GLASMA_ERR_ARG;
    Host h;
    int rc = alloc(ctx, h);
    if (rc) return rc;
    // start from the current fields, so that a file left out (NULL) keeps its field
    if ((rc = glasma_get_fields(ctx, h.U.data(), h.E.data(), h.Jp.data(), h.Jm.data(), h.JpS.data(), h.JmS.data())))
        return rc;
    if (ufile && (rc = load_links(ctx, ufile, 18, h.U))) return rc;
    if (efile && (rc = load_links(ctx, efile, 8, h.E))) return rc;
    if (jfile && (rc = load_currents(ctx, jfile, h))) return rc;
    if ((rc = glasma_set_fields(ctx, h.U.data(), h.E.data(), h.Jp.data(), h.Jm.data(), h.JpS.data(), h.JmS.data())))
        return rc;
    return glasma_set_time(ctx, t, std::lround(t / ctx->p.dt));
}

int glasma_save_state(glasma_ctx* ctx, const char* file) {
    if (!ctx || !file) return GLASMA_ERR_ARG;
    Host h;
    int rc = alloc(ctx, h);
    if (rc) return rc;
    if ((rc = glasma_get_fields(ctx, h.U.data(), h.E.data(), h.Jp.data(), h.Jm.data(), h.JpS.data(), h.JmS.data())))
        return rc;
    if ((rc = glasma_get_uold(ctx, h.UOld.data()))) return rc;
    double t;
    long n;
    glasma_get_time(ctx, &t, &n);
    std::ofstream f(file, std::ios::binary);
    if (!f) return io_fail(ctx, std::string("cannot open ") + file + " for writing");
    const int32_t hdr[5] = {ctx->p.nc, ctx->p.n[0], ctx->p.n[1], ctx->p.n[2], ctx->p.bc_z};
    const int64_t steps = n;
    f.write(kMagic, sizeof(kMagic));
    f.write((const char*)hdr, sizeof(hdr));
    f.write((const char*)ctx->p.a, sizeof(ctx->p.a));
    f.write((const char*)&ctx->p.dt, sizeof(double));
    f.write((const char*)&t, sizeof(double));
    f.write((const char*)&steps, sizeof(steps));
    for (const std::vector<double>* v : {&h.U, &h.UOld, &h.E, &h.Jp, &h.Jm, &h.JpS, &h.JmS})
        f.write((const char*)v->data(), (std::streamsize)(v->size() * sizeof(double)));
    if (!f) return io_fail(ctx, std::string("write error in ") + file);
    return 0;
}

int glasma_load_state(glasma_ctx* ctx, const char* file) {
    if (!ctx || !file) return GLASMA_ERR_ARG;
    std::ifstream f(file, std::ios::binary);
    if (!f) return io_fail(ctx, std::string("cannot open ") + file);
    char magic[8];
    int32_t hdr[5];
    double a[3], dt, t;
    int64_t steps;
    f.read(magic, sizeof(magic));
    f.read((char*)hdr, sizeof(hdr));
    f.read((char*)a, sizeof(a));
    f.read((char*)&dt, sizeof(dt));
    f.read((char*)&t, sizeof(t));
    f.read((char*)&steps, sizeof(steps));
    if (!f || std::memcmp(magic, kMagic, sizeof(kMagic)) != 0) return io_fail(ctx, std::string(file) + " is not a state file");
    if (hdr[0] != ctx->p.nc || hdr[1] != ctx->p.n[0] || hdr[2] != ctx->p.n[1] || hdr[3] != ctx->p.n[2] ||
        hdr[4] != ctx->p.bc_z || a[0] != ctx->p.a[0] || a[1] != ctx->p.a[1] || a[2] != ctx->p.a[2] || dt != ctx->p.dt)
        return io_fail(ctx, std::string(file) + " holds another lattice, spacing, time step or boundary");
    Host h;
    int rc = alloc(ctx, h);
    if (rc) return rc;
    for (std::vector<double>* v : {&h.U, &h.UOld, &h.E, &h.Jp, &h.Jm, &h.JpS, &h.JmS})
        f.read((char*)v->data(), (std::streamsize)(v->size() * sizeof(double)));
    if (!f) return io_fail(ctx, std::string(file) + " is truncated");
    if ((rc = glasma_set_fields(ctx, h.U.data(), h.E.data(), h.Jp.data(), h.Jm.data(), h.JpS.data(), h.JmS.data())))
        return rc;
    if ((rc = glasma_set_uold(ctx, h.UOld.data()))) return rc;
    return glasma_set_time(ctx, t, (long)steps);
}

}  // extern "C"
