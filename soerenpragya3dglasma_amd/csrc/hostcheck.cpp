// Host (g++) build of the per-site product code in su3.h / site.h, for CPU tests of the product
// arithmetic against the oracle and scipy.  Not part of libglasma and never a fallback for it: the
// device library is the only execution path of the engine.  Field arrays use the device's tiled layout.
#include <cstring>

#include "site.h"

using namespace glasma;

static void to_m3(const double* in, M3& A) {  // interleaved complex column-major -> M3
    for (int k = 0; k < 9; k++) { A.re[k] = in[2 * k]; A.im[k] = in[2 * k + 1]; }
}
static void from_m3(const M3& A, double* out) {
    for (int k = 0; k < 9; k++) { out[2 * k] = A.re[k]; out[2 * k + 1] = A.im[k]; }
}

extern "C" {

void hc_su3_expi(double c, const double* al, double* X) { M3 m; expi(c, al, m); from_m3(m, X); }
void hc_su3_expi_m1(double c, const double* al, double* X) { M3 m; expi_m1(c, al, m); from_m3(m, X); }
void hc_su3_ilog(double c, const double* U, double* out) { M3 m; to_m3(U, m); ilog(c, m, out); }
void hc_su3_adj_mul(const double* V, const double* al, double* out) { M3 m; to_m3(V, m); adj_mul(m, al, out); }
void hc_su3_inv_adj_mul(const double* V, const double* al, double* out) { M3 m; to_m3(V, m); inv_adj_mul(m, al, out); }
void hc_su3_retrigenu(const double* V, double* out) { M3 m; to_m3(V, m); retr_igen(m, out); }

// One fused leapfrog step over tiled device-layout arrays, sequentially over sites.
void hc_step(const int* n, const double* a, int bc, double dt, const double* U, double* Un, double* E,
             const double* J, double* Jn, double* UOld) {
    StepArgs A;
    A.U = U; A.Un = Un; A.E = E; A.J = J; A.Jn = Jn; A.UOld = UOld;
    A.g.nx = n[0]; A.g.ny = n[1]; A.g.nz = n[2]; A.g.bc = bc;
    A.g.P = (long)n[0] * n[1];
    A.g.S = A.g.P * n[2];
    fill_step_coefficients(A, a, dt);
    for (long s = 0; s < A.g.S; s++) step_site<unsigned long>(A, (unsigned long)s);
}

}
