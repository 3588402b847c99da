// sanitize_main — runs every extern "C" entry point of the oracle (cpu_ref.cpp, compiled into this
// translation unit) on small lattices, for a build with AddressSanitizer and UndefinedBehaviorSanitizer
// (`make -C oracle sanitize`).  TEST INFRASTRUCTURE ONLY, like the oracle itself.  Exit status 0 means
// every call ran without a sanitizer report (reports abort the program: -fno-sanitize-recover=all).
#include "cpu_ref.cpp"

#include <cstdio>
#include <random>

namespace {

struct State {
    int N[3];
    long S;  // sites with ghosts
    std::vector<double> U, UOld, E, Jp, Jm, JpS, JmS;
    State(int nx, int ny, int nz, int nc, std::mt19937_64& rng) : N{nx, ny, nz} {
        S = (long)nx * ny * (nz + 2);
        const int ms = nc == 3 ? 18 : 4, ng = nc == 3 ? 8 : 3;
        U.assign(3 * ms * S, 0.0);
        std::normal_distribution<double> nd(0.0, 0.3);
        for (long s = 0; s < 3 * S; s++) {
            double* u = &U[ms * s];
            if (nc == 3) {
                double al[8];
                for (double& v : al) v = nd(rng);
                ref_su3_expi(1.0, al, u);
            } else {
                double al[3];
                for (double& v : al) v = nd(rng);
                ref_su2_expi(1.0, al, u);
            }
        }
        UOld = U;
        E.assign(3 * ng * S, 0.0);
        for (double& v : E) v = nd(rng);
        for (auto* J : {&Jp, &Jm, &JpS, &JmS}) {
            J->assign(ng * S, 0.0);
            for (double& v : *J) v = nd(rng);
        }
    }
};

}  // namespace

int main() {
    std::mt19937_64 rng(5);
    const double a[3] = {1.0, 0.9, 0.8}, a1[3] = {1.0, 1.0, 1.0};
    ref_set_threads(2);
    // SU(3) algebra
    {
        double al[8] = {0.1, -0.2, 0.3, 0.05, -0.4, 0.2, 0.15, -0.1}, X[18], Y[18], C[18], v[8], adj[64];
        ref_su3_expi(0.7, al, X);
        ref_su3_expi(-0.3, al, Y);
        for (int m = 0; m < 4; m++) ref_su3_mul(X, Y, C, m);
        ref_su3_retrigenu(C, v);
        (void)ref_su3_retridminusu(C);
        (void)ref_su3_unitarity_norm(C);
        ref_su3_ilog(0.7, X, v);
        ref_su3_adj_mul(X, al, v);
        ref_su3_inv_adj_mul(X, al, v);
        ref_su3_get_adjoint(X, adj);
        ref_max_trace_projection(C, Y);
        double s2[4], b[3] = {0.2, -0.1, 0.3};
        ref_su2_expi(0.5, b, s2);
        ref_su2_ilog(0.5, s2, b);
    }
    // leapfrog and observables, SU(3) and SU(2), open and periodic z
    for (int nc : {3, 2})
        for (int bc : {0, 1}) {
            State f(5, 4, 6, nc, rng);
            double out[9], pz[6], mx, rms, emt[60], prof[24], cur[32], tmap[40];
            ref_update(nc, f.N, a, bc, 0.1, 2, f.U.data(), f.UOld.data(), f.E.data(), f.Jp.data(), f.Jm.data(),
                       f.JpS.data(), f.JmS.data());
            ref_sync_ghosts(nc, f.N, bc, f.E.data(), 3L * (nc == 3 ? 8 : 3));
            ref_bulk(nc, f.N, a, bc, f.U.data(), f.E.data(), f.Jp.data(), f.Jm.data(), out);
            ref_gauss(nc, f.N, a, bc, f.U.data(), f.E.data(), f.JpS.data(), f.JmS.data(), pz, &mx, &rms);
            ref_unitarity(nc, f.N, a, bc, f.U.data(), &mx, &rms);
            ref_emt(nc, f.N, a, bc, f.U.data(), f.UOld.data(), f.E.data(), emt);
            ref_field_profile(nc, f.N, a, bc, f.U.data(), f.E.data(), prof);
            ref_current_profile(nc, f.N, f.Jp.data(), f.Jm.data(), f.JpS.data(), f.JmS.data(), cur);
            ref_transverse_map(nc, f.N, a, bc, f.U.data(), f.E.data(), tmap);
        }
    // FFT and RNG
    {
        std::vector<double> d(2 * 8 * 6 * 5 * 7);
        for (double& v : d) v = 0.25;
        const int dims[3] = {6, 5, 7};
        ref_fft_nd(d.data(), 3, dims, 8, -1);
        ref_fft_nd(d.data(), 2, dims, 8 * 7, +1);
        std::vector<uint32_t> raw(100);
        ref_mt_raw(7, 0, 100, raw.data());
        ref_mt_raw(7, 1, 100, raw.data());
        std::vector<double> g(100);
        ref_gauss_draws(7, 100, g.data());
    }
    // MV initial conditions, the sheet from a charge, gauge fixing, Gauss restoration
    {
        const int N[3] = {6, 5, 8};
        std::vector<double> xl(8 * 30), xr(8 * 30);
        ref_draw_charges(3, 6, 5, 0.5, 0.5, xl.data(), xr.data());
        for (int driver = 0; driver < 3; driver++) {
            State f(6, 5, 8, 3, rng);
            ref_mv_initial_state(driver, N, a, 0, 0.25, 1.65, 0.2, 2.0, 2.0, 1.0, 1.0, 2, 5, xl.data(), xr.data(),
                                 f.U.data(), f.UOld.data(), f.E.data(), f.Jp.data(), f.Jm.data(), f.JpS.data(),
                                 f.JmS.data());
        }
        State f(6, 5, 8, 3, rng);
        std::vector<double> rho(8 * f.S, 0.01);
        ref_left_sheet_from_charge(N, a, 0, 0.25, 1.65, 0.2, rho.data(), f.U.data(), f.UOld.data(), f.E.data(),
                                   f.Jp.data(), f.JpS.data());
        std::vector<double> G(18 * f.S, 0.0), UG(f.U.size(), 0.0), EG(f.E.size(), 0.0);
        for (int algo = 0; algo < 3; algo++) {
            std::fill(G.begin(), G.end(), 0.0);
            for (long s = 0; s < f.S; s++) G[18 * s] = G[18 * s + 8] = G[18 * s + 16] = 1.0;
            UG = f.U;
            double st[7];
            ref_coulomb_gauge_algo(N, a1, f.U.data(), f.E.data(), G.data(), UG.data(), EG.data(), algo, 1e-12, 4,
                                   algo == 0 ? 2.0 : (algo == 1 ? 0.3 : 0.0), st);
        }
        double st[7];
        ref_coulomb_gauge(N, a1, f.U.data(), f.E.data(), G.data(), UG.data(), EG.data(), 1e-12, 3, 2.0, st);
        ref_gauge_transform(N, a1, 0, G.data(), f.U.data(), f.E.data(), f.Jp.data(), f.Jm.data(), f.JpS.data(),
                            f.JmS.data());
        for (int variant = 0; variant < 2; variant++) {
            double gs[3];
            ref_gauss_restore(N, a1, 0, f.U.data(), f.E.data(), f.JpS.data(), f.JmS.data(), variant, 1e-12, 5, gs);
        }
    }
    // physical-charge nuclei
    {
        const int N[3] = {6, 5, 12};
        const long n = 8L * 6 * 5 * 12;
        std::vector<double> nl(n), nr(n), pl(3 * 16), pr(3 * 16), CF(n);
        if (ref_tmd_draw(4, N, 16, nl.data(), nr.data(), pl.data(), pr.data()) != 0) return 2;
        const double prm[6] = {0.33, 0.33, 500 / 0.197, 0.13, 0.13, 7.7e-4};
        const int ip[3] = {16, 3, 8};
        ref_charge_fluctuation(N, a1, prm, ip, nl.data(), CF.data());
        (void)ref_nucleus_profile(prm, ip, pl.data(), 0.1, 0.2, 0.0);
        for (int driver = 0; driver < 3; driver++) {
            State f(6, 5, 12, 3, rng);
            std::vector<double> rl(8 * f.S), rr(8 * f.S);
            ref_tmd_initial_state(driver, N, a1, 0, 0.25, prm, ip, nl.data(), nr.data(), pl.data(), pr.data(),
                                  f.U.data(), f.UOld.data(), f.E.data(), f.Jp.data(), f.Jm.data(), f.JpS.data(),
                                  f.JmS.data(), rl.data(), rr.data());
        }
    }
    std::printf("sanitize: every oracle entry point ran\n");
    return 0;
}
