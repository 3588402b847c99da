// The fused FP64 leapfrog kernel: one launch per Dynamics::SUNc::Update (SC/DYNAMICS/SUNcDynamics.cpp:353-378).
//
// A block of 256 threads advances one tile of 64 sites (step_tile4): wavefronts 0, 1, 2 own link
// direction 0, 1, 2 of every site (link_update in site.h), wavefront 3 forms J+-stat'(x) and
// J+-stat'(x+z) and hands them to wavefront 2 through LDS, which applies the J advection with its own
// half-step link.  Splitting a site over wavefronts keeps each wavefront's live state to a few matrices.
// When a tile is made of whole lattice rows (nx a multiple or a divisor of 64), the link operands on those
// rows (own links and their x neighbours, 18 of the 41 matrix loads per site) come from an LDS copy made
// once per block; that form runs three wavefronts per SIMD, the plain form two.
//
// HBM traffic per site-update: read U (432 B) + E (192 B) + J (256 B), write the same: 1760 B, the
// compulsory minimum.  The neighbour link matrices are re-read through L2 / MALL, so the order in
// which the 8 XCDs (each with its own 4 MB L2) visit the lattice matters: workgroups are dealt to the
// XCDs round-robin, and with the column-band orders (TileMap mode 2 in site.h) each XCD walks its bands
// through z, so the y and z neighbours of a tile are read by the same XCD shortly before or after it (L2
// hits).  Measured at 256 x 256 x 1024 with the tuned 1 x 16 bands and three waves per SIMD, reads come to
// 1.40x the compulsory bytes (the y-halo rows of a band belong to another XCD, and the tiles in flight
// outgrow an XCD's 4 MB L2); writes are exactly compulsory.
//
// GLASMA_STEP_WAVES (2 or 3: register budget of 256 or 168 per lane), GLASMA_STEP_FOUR=2 (plain form only) and
// GLASMA_STEP_MAP / _BW / _BH select variants for measurement sweeps; the defaults are the fastest
// measured on MI355X.
#include <cstdlib>

#include "internal.h"

namespace glasma {

static int env_int(const char* name, int dflt) {
    const char* v = std::getenv(name);
    return v ? std::atoi(v) : dflt;
}


// The links of a tile's sites in LDS, slot 1 + lane, when row_tiles holds: nx a divisor of 64 (64 / nx whole
// rows, whose x neighbours all lie in the tile), or the tile is one row segment or two (the end of row r and the
// start of row r + 1).  The x neighbours past a segment's ends (x periodic within each row) sit in halo slots:
// slot 0 left of lane 0, slot 65 right of lane 63, and with two segments slot 66 right of the first segment's
// last lane (x = nx - 1 -> 0 of row r) and slot 67 left of the second segment's first lane (x = 0 -> nx - 1 of
// row r + 1).
constexpr int kRow = kTile + 2;     // one segment or whole rows: halo slots 0 and 65
constexpr int kRowSeg = kTile + 4;  // two segments: halo slots 0, 65, 66, 67

// The LDS slots of the x + 1 and x - 1 neighbours of this lane's site, and how many halo slots the tile uses (0;
// 2: slots 0, 65; 4: slots 0, 65, 66, 67).
struct RowSlots {
    int sp, sm, nhalo;
};
// x0: the x of the tile's first site (lane 0's x)
__device__ __forceinline__ RowSlots row_slots(const Geom& g, int x0, int lane, int x) {
    RowSlots r;
    const int nx = g.nx;
    if (kTile % nx == 0) {  // whole rows in the tile
        r.sp = lane + 1 + (x + 1 == nx ? 1 - nx : 1);
        r.sm = lane + 1 + (x == 0 ? nx - 1 : -1);
        r.nhalo = 0;
        return r;
    }
    const int k = nx - x0;  // the tile's sites in its first row
    if (k >= kTile) {                      // one segment
        r.sp = lane + 2;
        r.sm = lane;
        r.nhalo = 2;
        return r;
    }
    r.sp = lane == k - 1 ? kTile + 2 : (lane == kTile - 1 ? kTile + 1 : lane + 2);
    r.sm = lane == 0 ? 0 : (lane == k ? kTile + 3 : lane);
    r.nhalo = 4;
    return r;
}
// the site held by halo slot h (0: slot 0, 1: slot 65, 2: slot 66, 3: slot 67) of the tile starting at `first`,
// whose first site has x = x0
__device__ __forceinline__ long halo_site(const Geom& g, long first, int x0, int h) {
    const int nx = g.nx;
    const long r1 = first - x0;  // start of the tile's first row
    if (h == 0) return r1 + (x0 == 0 ? nx - 1 : x0 - 1);
    const int k = nx - x0;
    if (h == 1) {
        if (k >= kTile) return r1 + (x0 + kTile == nx ? 0 : x0 + kTile);
        const int x2 = kTile - k;  // right of lane 63, at x = kTile - k - 1 of row r + 1
        return r1 + nx + (x2 == nx ? 0 : x2);
    }
    return h == 2 ? r1 : r1 + 2L * nx - 1;
}

// LDS of one block of the four-wavefront step (RL slots per row: kRow or kRowSeg).  The row length is a template
// parameter so that the whole-row form keeps the 66-slot layout, whose LDS offsets let it fit 168 registers
// without spills (68 slots cost two spilled doubles).
template <int RL>
struct StepShared {
    double row[3][18][RL];  // U_mu on the tile's row(s) and halo sites, plane pairs as in ldm (re, im of element k)
    double j[6][8][kTile];  // J+stat'(x), J-stat'(x), J+stat'(x+z), J-stat'(x+z), J+(x), J-(x)
    int ready;
};

template <int RL>
__device__ __forceinline__ void ld_row(const double (*r)[RL], int slot, M3& M) {
#pragma unroll
    for (int k = 0; k < 9; k++) {
        M.re[k] = r[2 * k][slot];
        M.im[k] = r[2 * k + 1][slot];
    }
}

// The two staples of link MU in transverse direction NU (link_update's loop body) with the operands on
// the tile's own rows -- U_nu(x), U_nu(x+x) for MU = x, U_mu(x+-x) and U_x(x-x) for NU = x -- read from
// the LDS copy instead of global memory (slots: own 1 + lane, x+1 at sp, x-1 at sm).  Same products in the
// same order as link_update.
// PRE: the forward staple's operands from global memory, U_nu(x+mu) (MU != 0) and U_mu(x+nu) (NU != 0), come in
// pre[0] and pre[1] (read before the row copy's barrier, first_staple_operands).
template <class I, int MU, int NU, int RL, bool PRE = false>
__device__ __forceinline__ void staples_row(const StepArgs& A, I s, int x, int y, int z, int lane, int sp, int sm,
                                            const double (*row)[18][RL], M3& Sig, const M3* pre = nullptr) {
    const Geom& g = A.g;
    const double cn = A.cE[NU];
    M3 P1, P2, T;
    // forward staple U_nu(x+mu) U_mu^dag(x+nu) U_nu^dag(x)
    if (MU == 0) ld_row(row[NU], sp, P1);
    else if (PRE) P1 = pre[0];
    else ldm<I>(A.U, kUPlanes, 18 * NU, nbr<I>(g, x, y, z, MU, 1, NU, 0), P1);
    if (NU == 0) ld_row(row[MU], sp, P2);
    else if (PRE) P2 = pre[1];
    else ldm<I>(A.U, kUPlanes, 18 * MU, nbr<I>(g, x, y, z, NU, 1, MU, 0), P2);
    mul_nd(P1, P2, T);
    ld_row(row[NU], lane + 1, P1);
    acc_mul_nd(cn, T, P1, Sig);
    // backward staple U_nu^dag(x+mu-nu) U_mu^dag(x-nu) U_nu(x-nu)
    ldm<I>(A.U, kUPlanes, 18 * NU, nbr<I>(g, x, y, z, MU, 1, NU, -1), P1);
    if (NU == 0) {
        ld_row(row[MU], sm, P2);
    } else {
        ldm<I>(A.U, kUPlanes, 18 * MU, nbr<I>(g, x, y, z, NU, -1, MU, 0), P2);
    }
    mul_dd(P1, P2, T);
    if (NU == 0) ld_row(row[0], sm, P1);
    else ldm<I>(A.U, kUPlanes, 18 * NU, nbr<I>(g, x, y, z, NU, -1, MU, 0), P1);
    acc_mul_nn(cn, T, P1, Sig);
}

// wait for the J wavefront's hand-over (set unconditionally once its LDS stores are complete).  The wait is
// capped so that no wave can spin forever; false if the cap was reached, and the caller then writes NaN
// currents (glasma_bulk reports non-finite sums as GLASMA_ERR_NUMERIC) instead of silently using stale values.
template <class SH>
__device__ __forceinline__ bool wait_ready(SH& sh) {
    bool ok = false;
    for (int spin = 0; spin < (1 << 24); spin++) {
        if (__hip_atomic_load(&sh.ready, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) {
            ok = true;
            break;
        }
        __builtin_amdgcn_s_sleep(2);
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    return ok;
}

// the global operands of link MU's first forward staple (NU = MU + 1), read before the row copy's barrier so that
// their latency overlaps the barrier's wait: pre[0] = U_nu(x+mu) when MU != 0, pre[1] = U_mu(x+nu) when NU != 0
template <class I, int MU>
__device__ __forceinline__ void first_staple_operands(const StepArgs& A, int x, int y, int z, M3* pre) {
    constexpr int NU = (MU + 1) % 3;
    if (MU != 0) ldm<I>(A.U, kUPlanes, 18 * NU, nbr<I>(A.g, x, y, z, MU, 1, NU, 0), pre[0]);
    if (NU != 0) ldm<I>(A.U, kUPlanes, 18 * MU, nbr<I>(A.g, x, y, z, NU, 1, MU, 0), pre[1]);
}

// link_update with the row operands from LDS (PRE: the first staple's global operands in pre, first_staple_operands)
template <class I, int MU, int RL, bool PRE = false>
__device__ __forceinline__ void link_update_row(const StepArgs& A, I s, int x, int y, int z, int lane, int sp,
                                                int sm, const double (*row)[18][RL], M3& Uh,
                                                const M3* pre = nullptr) {
    M3 Sig;
#pragma unroll
    for (int k = 0; k < 9; k++) { Sig.re[k] = 0.0; Sig.im[k] = 0.0; }
    // scheduling barriers keep the compiler from hoisting the next staple pair's loads into this one:
    // that fits the form into 168 registers without spills (three waves per SIMD, 31.3 -> 30.1 ms at
    // 256 x 256 x 1024), the waves hiding the latency instead
    staples_row<I, MU, (MU + 1) % 3, RL, PRE>(A, s, x, y, z, lane, sp, sm, row, Sig, pre);
    __builtin_amdgcn_sched_barrier(0);
    staples_row<I, MU, (MU + 2) % 3, RL>(A, s, x, y, z, lane, sp, sm, row, Sig);
    __builtin_amdgcn_sched_barrier(0);
    M3 Uo;
    ld_row(row[MU], lane + 1, Uo);
    link_finish<I, MU>(A, s, Uo, Sig, Uh);
}

// Four-wavefront step of the tile starting at site `first`.  Waves 0-2 update the link directions; wave 3
// forms J+-stat'(x) and J+-stat'(x+z) (the cheaper work) and hands them, with the J+-(x) it read, to
// wave 2 through LDS.  Wave 2 applies the J advection with its own U_h,z as soon as its link is done, so
// the block ends with its slowest link wavefront instead of a serial advection after a block barrier.
// The hand-over is a workgroup-scope flag: wave 3 sets it unconditionally once its stores to LDS are
// complete, so wave 2's wait always ends; the wait is capped all the same.
//
// ROW (row_tiles: whole rows, or one or two row segments): waves 0-2 first copy U_w of the tile and of its halo
// sites (row_slots) into LDS, and the 18 of 41 link operands per site that lie on those rows are read from
// there.
// This takes about a third of the vector-memory accesses off the L1, whose pipeline is what the
// kernel waits on.
template <class I, bool ROW, bool SEG, int RL = SEG ? kRowSeg : kRow>
__device__ __forceinline__ void step_tile4(const StepArgs& A, long first, StepShared<RL>& sh) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const long sl = first + lane;
    const bool valid = sl < A.g.S;
    const I s = (I)(valid ? sl : 0);
    int x, y, z;
    site_coords<I>(A.g, s, x, y, z);
    // LDS slots of x+1 and x-1: the next / previous lane, or halo slots; in a tile of several short rows, the
    // lane of the periodic neighbour in the same row; with two row segments (SEG), row_slots
    int sp, sm, nhalo, x0 = 0;
    const bool one_row = A.g.nx % kTile == 0;
    if (SEG) {
        x0 = __builtin_amdgcn_readfirstlane(x);  // lane 0's x: every lane of the wave is active here
        const RowSlots rs = row_slots(A.g, x0, lane, x);
        sp = rs.sp;
        sm = rs.sm;
        nhalo = rs.nhalo;
    } else {
        sp = one_row ? lane + 2 : lane + 1 + (x + 1 == A.g.nx ? 1 - A.g.nx : 1);
        sm = one_row ? lane : lane + 1 + (x == 0 ? A.g.nx - 1 : -1);
        nhalo = one_row ? 2 : 0;
    }
    if (ROW && wave < 3) {
        // lanes past the end of the lattice (last tile only, whole-row tiles only) copy site 0; no valid lane
        // reads their slots
        M3 M;
        ldm<I>(A.U, kUPlanes, 18 * wave, s, M);
#pragma unroll
        for (int k = 0; k < 9; k++) {
            sh.row[wave][2 * k][lane + 1] = M.re[k];
            sh.row[wave][2 * k + 1][lane + 1] = M.im[k];
        }
        if (lane < nhalo) {
            long hs;
            if (SEG) {
                hs = halo_site(A.g, first, x0, lane);
            } else {
                const int xt = x - lane;  // the tile's first x
                hs = (long)s - x + (lane == 0 ? xwrap(xt - 1, A.g.nx) : xwrap(xt + kTile, A.g.nx));
            }
            ldm<I>(A.U, kUPlanes, 18 * wave, (I)hs, M);
            const int slot = lane == 0 ? 0 : kTile + lane;
#pragma unroll
            for (int k = 0; k < 9; k++) {
                sh.row[wave][2 * k][slot] = M.re[k];
                sh.row[wave][2 * k + 1][slot] = M.im[k];
            }
        }
    }
    if (threadIdx.x == 0) sh.ready = 0;
    __syncthreads();
    if (wave == 0) {
        M3 Uh;
        if (ROW) {
            if (valid) link_update_row<I, 0, RL>(A, s, x, y, z, lane, sp, sm, sh.row, Uh);
        } else if (valid) {
            link_update<I, 0>(A, s, x, y, z, Uh);
        }
    } else if (wave == 1) {
        M3 Uh;
        if (ROW) {
            if (valid) link_update_row<I, 1, RL>(A, s, x, y, z, lane, sp, sm, sh.row, Uh);
        } else if (valid) {
            link_update<I, 1>(A, s, x, y, z, Uh);
        }
    } else if (wave == 3) {
        if (valid) {
            double jps[8], jms[8], jp[8], jm[8];
            const int zm = zdn(A.g, z);
            jstat_new_keep<I>(A, s, (I)(sl + A.g.P * (long)(zm - z)), jps, jms, jp, jm);
            stv<I>(A.Jn, kJPlanes, JPS * 8, s, jps);
            stv<I>(A.Jn, kJPlanes, JMS * 8, s, jms);
#pragma unroll
            for (int a = 0; a < 8; a++) {
                sh.j[0][a][lane] = jps[a];
                sh.j[1][a][lane] = jms[a];
                sh.j[4][a][lane] = jp[a];
                sh.j[5][a][lane] = jm[a];
            }
            if (ROW) {
                // U_z(x) of J+-stat'(x+z) from the row copy; on the open top layer J+-stat'(x+z) is
                // J+-stat'(x) itself (jstat_new_up), already in jps, jms
                const int zp = zup(A.g, z);
                if (zp != z) {
                    M3 Uz;
                    ld_row(sh.row[2], lane + 1, Uz);
                    jstat_core<I>(A, (I)(sl + A.g.P * (long)(zp - z)), s, Uz, jps, jms);
                }
            } else {
                jstat_new_up<I>(A, s, z, jps, jms);
            }
#pragma unroll
            for (int a = 0; a < 8; a++) {
                sh.j[2][a][lane] = jps[a];
                sh.j[3][a][lane] = jms[a];
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        if (lane == 0) __hip_atomic_store(&sh.ready, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    } else {
        M3 Uh;
        if (ROW) {
            if (valid) link_update_row<I, 2, RL>(A, s, x, y, z, lane, sp, sm, sh.row, Uh);
        } else if (valid) {
            link_update<I, 2>(A, s, x, y, z, Uh);
        }
        const bool handed_over = wait_ready(sh);
        if (valid) {
            double jps[8], jms[8], jpsU[8], jmsU[8], jp[8], jm[8];
#pragma unroll
            for (int a = 0; a < 8; a++) {
                jps[a] = sh.j[0][a][lane];
                jms[a] = sh.j[1][a][lane];
                jpsU[a] = sh.j[2][a][lane];
                jmsU[a] = sh.j[3][a][lane];
                jp[a] = sh.j[4][a][lane];
                jm[a] = sh.j[5][a][lane];
            }
            if (!handed_over) jp[0] = jm[0] = __builtin_nan("");
            j_advect_with<I>(A, s, Uh, jps, jms, jpsU, jmsU, jp, jm);
        }
    }
}

// One link wavefront of step_tile4_pre (link direction MU): the row copy of U_MU (ROW) with the first staple's global
// operands read in the same burst, the block barrier, the link update and, for MU = 2, the J advection after the J
// wavefront's hand-over.  Each wavefront's branch holds its own barrier (the branches are uniform: every wavefront
// runs exactly one), so that the operands read before it stay live in that branch only.
template <class I, int MU, bool ROW, bool SEG, int RL>
__device__ __forceinline__ void link_wave(const StepArgs& A, long first, I s, bool valid, int x, int y, int z, int lane,
                                          int sp, int sm, int nhalo, int x0, StepShared<RL>& sh) {
    M3 pre[2];
    if (ROW) {
        // lanes past the end of the lattice (last tile only, whole-row tiles only) copy site 0; no valid lane
        // reads their slots
        M3 M;
        ldm<I>(A.U, kUPlanes, 18 * MU, s, M);
        first_staple_operands<I, MU>(A, x, y, z, pre);
#pragma unroll
        for (int k = 0; k < 9; k++) {
            sh.row[MU][2 * k][lane + 1] = M.re[k];
            sh.row[MU][2 * k + 1][lane + 1] = M.im[k];
        }
        if (lane < nhalo) {
            long hs;
            if (SEG) {
                hs = halo_site(A.g, first, x0, lane);
            } else {
                const int xt = x - lane;  // the tile's first x
                hs = (long)s - x + (lane == 0 ? xwrap(xt - 1, A.g.nx) : xwrap(xt + kTile, A.g.nx));
            }
            ldm<I>(A.U, kUPlanes, 18 * MU, (I)hs, M);
            const int slot = lane == 0 ? 0 : kTile + lane;
#pragma unroll
            for (int k = 0; k < 9; k++) {
                sh.row[MU][2 * k][slot] = M.re[k];
                sh.row[MU][2 * k + 1][slot] = M.im[k];
            }
        }
    }
    if (MU == 0 && lane == 0) sh.ready = 0;
    __syncthreads();
    M3 Uh;
    if (ROW) {
        if (valid) link_update_row<I, MU, RL, true>(A, s, x, y, z, lane, sp, sm, sh.row, Uh, pre);
    } else if (valid) {
        link_update<I, MU>(A, s, x, y, z, Uh);
    }
    if (MU == 2) {
        const bool handed_over = wait_ready(sh);
        if (valid) {
            double jps[8], jms[8], jpsU[8], jmsU[8], jp[8], jm[8];
#pragma unroll
            for (int a = 0; a < 8; a++) {
                jps[a] = sh.j[0][a][lane];
                jms[a] = sh.j[1][a][lane];
                jpsU[a] = sh.j[2][a][lane];
                jmsU[a] = sh.j[3][a][lane];
                jp[a] = sh.j[4][a][lane];
                jm[a] = sh.j[5][a][lane];
            }
            if (!handed_over) jp[0] = jm[0] = __builtin_nan("");
            j_advect_with<I>(A, s, Uh, jps, jms, jpsU, jmsU, jp, jm);
        }
    }
}

// step_tile4 with each link wavefront's branch holding its own barrier and the first staple's global operands read
// before it (link_wave).  Measured for the two-segment row form at 96 x 96 x 384: 1.69-1.72 against 1.71-1.75 ms; for
// whole rows at 256 x 256 x 1024 it was slower (29.96 against 29.62 ms: 4 spilled registers), so k_step4 keeps
// step_tile4 and the ensemble takes this form for its two-segment blocks.
template <class I, bool ROW, bool SEG, int RL = SEG ? kRowSeg : kRow>
__device__ __forceinline__ void step_tile4_pre(const StepArgs& A, long first, StepShared<RL>& sh) {
    // the wave index through readfirstlane: uniform to the compiler, so the per-wave branches are scalar branches
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const long sl = first + lane;
    const bool valid = sl < A.g.S;
    const I s = (I)(valid ? sl : 0);
    int x, y, z;
    site_coords<I>(A.g, s, x, y, z);
    if (wave == 3) {
        __syncthreads();
        if (valid) {
            double jps[8], jms[8], jp[8], jm[8];
            const int zm = zdn(A.g, z);
            jstat_new_keep<I>(A, s, (I)(sl + A.g.P * (long)(zm - z)), jps, jms, jp, jm);
            stv<I>(A.Jn, kJPlanes, JPS * 8, s, jps);
            stv<I>(A.Jn, kJPlanes, JMS * 8, s, jms);
#pragma unroll
            for (int a = 0; a < 8; a++) {
                sh.j[0][a][lane] = jps[a];
                sh.j[1][a][lane] = jms[a];
                sh.j[4][a][lane] = jp[a];
                sh.j[5][a][lane] = jm[a];
            }
            if (ROW) {
                // U_z(x) of J+-stat'(x+z) from the row copy; on the open top layer J+-stat'(x+z) is
                // J+-stat'(x) itself (jstat_new_up), already in jps, jms
                const int zp = zup(A.g, z);
                if (zp != z) {
                    M3 Uz;
                    ld_row(sh.row[2], lane + 1, Uz);
                    jstat_core<I>(A, (I)(sl + A.g.P * (long)(zp - z)), s, Uz, jps, jms);
                }
            } else {
                jstat_new_up<I>(A, s, z, jps, jms);
            }
#pragma unroll
            for (int a = 0; a < 8; a++) {
                sh.j[2][a][lane] = jps[a];
                sh.j[3][a][lane] = jms[a];
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        if (lane == 0) __hip_atomic_store(&sh.ready, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        return;
    }
    // LDS slots of x+1 and x-1: the next / previous lane, or halo slots; in a tile of several short rows, the
    // lane of the periodic neighbour in the same row; with two row segments (SEG), row_slots
    int sp, sm, nhalo, x0 = 0;
    const bool one_row = A.g.nx % kTile == 0;
    if (SEG) {
        x0 = __builtin_amdgcn_readfirstlane(x);  // lane 0's x: every lane of the wave is active here
        const RowSlots rs = row_slots(A.g, x0, lane, x);
        sp = rs.sp;
        sm = rs.sm;
        nhalo = rs.nhalo;
    } else {
        sp = one_row ? lane + 2 : lane + 1 + (x + 1 == A.g.nx ? 1 - A.g.nx : 1);
        sm = one_row ? lane : lane + 1 + (x == 0 ? A.g.nx - 1 : -1);
        nhalo = one_row ? 2 : 0;
    }
    if (wave == 0) link_wave<I, 0, ROW, SEG, RL>(A, first, s, valid, x, y, z, lane, sp, sm, nhalo, x0, sh);
    else if (wave == 1) link_wave<I, 1, ROW, SEG, RL>(A, first, s, valid, x, y, z, lane, sp, sm, nhalo, x0, sh);
    else link_wave<I, 2, ROW, SEG, RL>(A, first, s, valid, x, y, z, lane, sp, sm, nhalo, x0, sh);
}

// Four-wavefront variant (step_tile4); two such blocks fill the 8 wavefront slots of a CU at two waves
// per SIMD.
template <class I, int W, bool ROW, bool SEG>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(W, W))) k_step4(StepArgs A, TileMap tm) {
    __shared__ StepShared<SEG ? kRowSeg : kRow> sh;
    step_tile4<I, ROW, SEG>(A, tile_site(A.g, tm, blockIdx.x), sh);
}

// Ensemble step: one launch per form advances every event of a batch.  Block b works on the tile starting
// at site block_site[b] of event e = block_event[b] (each event's blocks in the order tile_site gives it
// alone); the event's arguments come from a device table (uniform per block, so they are read with scalar
// loads).  The per-site arithmetic is that of k_step4.  The blocks come in three groups by row_form: whole rows
// or one row segment per tile, then two segments (both the row-staged form at three waves per SIMD), then the
// plain form at two, as each runs alone.
template <int W, bool ROW, bool SEG>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(W, W)))
k_step_batch(const StepArgs* __restrict__ table, const int* __restrict__ block_event,
             const unsigned* __restrict__ block_site) {
    __shared__ StepShared<SEG ? kRowSeg : kRow> sh;
    const StepArgs A = table[block_event[blockIdx.x]];
    if constexpr (SEG) step_tile4_pre<unsigned int, ROW, SEG>(A, (long)block_site[blockIdx.x], sh);
    else step_tile4<unsigned int, ROW, SEG>(A, (long)block_site[blockIdx.x], sh);
}

int launch_step_batch(hipStream_t st, const StepArgs* table, const int* block_event, const unsigned* block_site,
                      long nrow, long nseg, long nplain) {
    if (nrow < 0 || nseg < 0 || nplain < 0 || nrow + nseg + nplain <= 0 || nrow > 0x7fffffffL ||
        nseg > 0x7fffffffL || nplain > 0x7fffffffL)
        return GLASMA_ERR_ARG;
    if (nrow > 0)
        hipLaunchKernelGGL((k_step_batch<3, true, false>), dim3((unsigned)nrow), dim3(256), 0, st, table, block_event,
                           block_site);
    if (nseg > 0)
        hipLaunchKernelGGL((k_step_batch<3, true, true>), dim3((unsigned)nseg), dim3(256), 0, st, table,
                           block_event + nrow, block_site + nrow);
    if (nplain > 0)
        hipLaunchKernelGGL((k_step_batch<2, false, false>), dim3((unsigned)nplain), dim3(256), 0, st, table,
                           block_event + nrow + nseg, block_site + nrow + nseg);
    return hipGetLastError() == hipSuccess ? 0 : GLASMA_ERR_HIP;
}

// the arguments of one step of a context (current buffers, optional UOld)
StepArgs step_args(const glasma_ctx* ctx, int cur, bool write_old) {
    StepArgs A;
    A.U = ctx->U[cur];
    A.Un = ctx->U[cur ^ 1];
    A.E = ctx->E;
    A.J = ctx->J[cur];
    A.Jn = ctx->J[cur ^ 1];
    A.UOld = write_old ? ctx->UOld : nullptr;
    A.g = ctx->g;
    fill_step_coefficients(A, ctx->p.a, ctx->p.dt);
    return A;
}

template <class I>
static void launch(int waves, int four, const TileMap& tm, const StepArgs& A, hipStream_t st) {
    dim3 grid((unsigned)(pad_sites(A.g.S) / kTile));
    // the row-staged form needs whole rows per tile (GLASMA_STEP_FOUR=2 forces the plain form); three
    // waves per SIMD (168 registers) measured faster for it (30.1 against 31.8 ms at 256 x 256 x 1024),
    // the plain form spills at three and runs at two
    const int form = four == 2 ? 0 : row_form(A.g);
    if (form == 2) {
        hipLaunchKernelGGL((k_step4<I, 3, true, true>), grid, dim3(256), 0, st, A, tm);
    } else if (form == 1) {
        if (waves == 2) hipLaunchKernelGGL((k_step4<I, 2, true, false>), grid, dim3(256), 0, st, A, tm);
        else hipLaunchKernelGGL((k_step4<I, 3, true, false>), grid, dim3(256), 0, st, A, tm);
    } else {
        if (waves == 3) hipLaunchKernelGGL((k_step4<I, 3, false, false>), grid, dim3(256), 0, st, A, tm);
        else hipLaunchKernelGGL((k_step4<I, 2, false, false>), grid, dim3(256), 0, st, A, tm);
    }
}

// tile order for a geometry: the requested mode when the lattice fits it, else the next simpler one
TileMap tile_map(const Geom& g, int mode, int bw, int bh) {
    TileMap m = {0, 0, 0, 0, 0, 0};
    if (mode == 2 && g.nx % kTile == 0 && bw > 0 && bh > 0) {
        const int tpr = g.nx / kTile;
        if (tpr % bw == 0 && g.ny % bh == 0) {
            const long bands = (long)(tpr / bw) * (g.ny / bh);
            if (bands % kXcds == 0) {
                m.mode = 2;
                m.bw = bw;
                m.bh = bh;
                m.nbx = tpr / bw;
                m.tpbp = bw * bh;
                return m;
            }
        }
    }
    if (mode == 3) {
        const long nt = pad_sites(g.S) / kTile;
        if (nt % kXcds == 0 && nt / kXcds <= 0x7fffffffL) {
            m.mode = 3;
            m.bpb = (int)(nt / kXcds);
            return m;
        }
    }
    if (mode >= 1 && g.P % (kXcds * kTile) == 0) {
        m.mode = 1;
        m.bpb = (int)(g.P / (kXcds * kTile));
    }
    return m;
}

TileMap default_tile_map(const Geom& g) {
    // GLASMA_STEP_MAP = mode (0, 1, 2) and, for mode 2, the band width in tiles and height in rows
    // (mode 2 with 4 x 4 bands measured fastest at 256 x 256 x 1024: 33.9 ms against 36.6 ms for mode 1)
    static const int map_mode = env_int("GLASMA_STEP_MAP", 2);
    static const int map_bw = env_int("GLASMA_STEP_BW", 4);
    static const int map_bh = env_int("GLASMA_STEP_BH", 4);
    return tile_map(g, map_mode, map_bw, map_bh);
}

TileMap ctx_tile_map(const glasma_ctx* ctx) { return ctx->tm_tuned ? ctx->tm : default_tile_map(ctx->g); }

// one step with explicit buffers and tile order (the context supplies the stream)
int launch_step_map(glasma_ctx* ctx, const StepArgs& A, const TileMap& tm, hipStream_t st) {
    static const int waves = env_int("GLASMA_STEP_WAVES", 0);  // 0: the measured best of each form
    static const int four = env_int("GLASMA_STEP_FOUR", 1);
    if (!st) st = ctx->stream;
    if (A.g.S < (1L << 32)) launch<unsigned int>(waves, four, tm, A, st);
    else launch<unsigned long>(waves, four, tm, A, st);
    return 0;
}

// one step with explicit buffers (the context supplies geometry, stream and tile order)
int launch_step_args(glasma_ctx* ctx, const StepArgs& A, hipStream_t st) {
    return launch_step_map(ctx, A, ctx_tile_map(ctx), st);
}

int launch_step(glasma_ctx* ctx, bool write_old) {
    const int rc = launch_step_args(ctx, step_args(ctx, ctx->cur, write_old));
    ctx->cur ^= 1;
    return rc;
}

}  // namespace glasma
