// Discrete adjoint behind include/fieldgrid_adjoint.h: dJ/d eps_r over a design box.
//
// No counterpart in the reference. One forward step (cpu.rs:267-307, the fused engine here)
// maps the state x = (E, H, psi) linearly: the H update (cpu.rs:48-107) and then the E update
// (cpu.rs:109-167), whose coefficients (ca, cb) are the only place eps enters. The adjoint state
// lam = (lamE, lamH, lamPsi) runs backward:
//   lam^n = M^T lam^{n+1} + dJ/dx^n,  dJ/dca_e += lamE_e^{n+1} E_e^n,  dJ/dcb_e += lamE_e^{n+1} curl_e^n.
// M^T is written in gather form, so each thread owns its outputs:
//   k_adj_e  transposes the E update: H-bar gathers the "a = cb lamE" terms of the E edges whose
//            curl reads it, at the node and its +1 neighbours (the shape of an H update); a is
//            kept as a field of its own (la), written by k_adj_h, so no coefficient is read here;
//   k_adj_h  transposes the H update and finishes lamE = ca lamE + (the "b = -k_mu lamH" terms
//            of the H components whose curl reads it, at the node and its -1 neighbours), and
//            advances every CPML psi adjoint into the other ping-pong copy.
// The forward pass keeps field snapshots every `every` steps; the backward pass restores one,
// recomputes that segment while taping (E^n, curl^n) on the design edges, then runs the adjoint
// over the segment. An f64 forward (k_h64 / k_e64) evaluates the same problem with the box's
// edge coefficients in f64, for finite-difference checks.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/fieldgrid_adjoint.h"
#include "host.h"

namespace fg {

// hv / hu adjoint terms of E edge C at node (i,j,k): with a = cb lamE (kept as its own field
// `la`, so no coefficient is looked up here), the E update's t1 = hv inv_dd_u + psi_u',
// t2 = hu inv_dd_v + psi_v', psi' = be psi + ce (h diff) raw_dd gives
//   hv-bar = a inv_dd_u + ce_u raw_dd_u (psiBar_u + a),  hu-bar = -a inv_dd_v + ce_v raw_dd_v (psiBar_v - a).
// Zero outside the grid; an edge with cb = 0 has a = 0 and a psi adjoint that stays 0.
template <int C>
__device__ __forceinline__ void e_terms(const Geo &g, const float *__restrict__ la, const float *__restrict__ lp,
                                        uint32_t i, uint32_t j, uint32_t k, float &hv, float &hu) {
    constexpr int U = (C + 1) % 3, V = (C + 2) % 3;
    hv = 0.f, hu = 0.f;
    if (i >= g.n0 || j >= g.n1 || k >= g.n2) return;
    const uint64_t id = (uint64_t(i) * g.n1 + j) * g.n2 + k;
    const float a = la[C * g.nn + id];
    const uint32_t pu = coord<U>(i, j, k), pv = coord<V>(i, j, k);
    hv = a * g.ax[U].inv_dd[pu];
    hu = -a * g.ax[V].inv_dd[pv];
    if (g.pml) {
        const int su = g.ax[U].slot_e[pu], sv = g.ax[V].slot_e[pv];
        if (su >= 0) hv += g.ax[U].ce[pu] * g.ax[U].raw_dd[pu] * (lp[psi_at<U>(g, 0, C, i, j, k, uint32_t(su))] + a);
        if (sv >= 0) hu += g.ax[V].ce[pv] * g.ax[V].raw_dd[pv] * (lp[psi_at<V>(g, 0, C, i, j, k, uint32_t(sv))] - a);
    }
}

// ev / eu adjoint terms of H component C at node (i,j,k): with b = -k_mu lamH, the H update's
// t1 = ev inv_d_u + psi_u', t2 = eu inv_d_v + psi_v', psi' = bh psi + ch (e diff) raw_d gives
//   ev-bar = b inv_d_u + ch_u raw_d_u (psiBar_u + b),  eu-bar = -b inv_d_v + ch_v raw_d_v (psiBar_v - b).
// Zero outside the grid and for components the H update skips.
template <int C>
__device__ __forceinline__ void h_terms(const Geo &g, const float *__restrict__ lamH, const float *__restrict__ lp,
                                        uint32_t i, uint32_t j, uint32_t k, float &ev, float &eu) {
    constexpr int U = (C + 1) % 3, V = (C + 2) % 3;
    ev = 0.f, eu = 0.f;
    if (i >= g.n0 || j >= g.n1 || k >= g.n2) return;
    const uint32_t pu = coord<U>(i, j, k), pv = coord<V>(i, j, k);
    if (pu + 1 >= dim<U>(g) || pv + 1 >= dim<V>(g)) return;
    const uint64_t id = (uint64_t(i) * g.n1 + j) * g.n2 + k;
    const float b = -g.k_mu * lamH[C * g.nn + id];
    ev = b * g.ax[U].inv_d[pu];
    eu = -b * g.ax[V].inv_d[pv];
    if (g.pml) {
        const int su = g.ax[U].slot_h[pu], sv = g.ax[V].slot_h[pv];
        if (su >= 0) ev += g.ax[U].ch[pu] * g.ax[U].raw_d[pu] * (lp[psi_at<U>(g, 1, C, i, j, k, uint32_t(su))] + b);
        if (sv >= 0) eu += g.ax[V].ch[pv] * g.ax[V].raw_d[pv] * (lp[psi_at<V>(g, 1, C, i, j, k, uint32_t(sv))] - b);
    }
}

// Design edges of one step's gradient, looked up per node of the box (nodes lo..hi per axis).
struct GradArgs {
    uint32_t lo[3], nb[3];     // first node and node counts of the box's edges
    const int32_t *map;        // [c][box node] -> design edge, -1 where eps does not enter
    const float2 *tape;        // this step's (E^n, curl^n) per design edge
    double *gca, *gcb;         // dJ/dca, dJ/dcb per design edge
};

// lamE^{n+1} (E^n, curl^n) summed into dJ/dca, dJ/dcb on the design edges at node (i,j,k)
__device__ __forceinline__ void grad_terms(const Geo &g, const GradArgs &ga, const float *__restrict__ lamE,
                                           uint32_t i, uint32_t j, uint32_t k, uint64_t id) {
    const uint32_t bi = i - ga.lo[0], bj = j - ga.lo[1], bk = k - ga.lo[2];
    if (bi >= ga.nb[0] || bj >= ga.nb[1] || bk >= ga.nb[2]) return;
    const uint64_t b = (uint64_t(bi) * ga.nb[1] + bj) * ga.nb[2] + bk, nb = uint64_t(ga.nb[0]) * ga.nb[1] * ga.nb[2];
    for (int c = 0; c < 3; c++) {
        const int32_t e = ga.map[c * nb + b];
        if (e < 0) continue;
        const double l = lamE[c * g.nn + id];
        const float2 tp = ga.tape[e];
        ga.gca[e] += l * double(tp.x);
        ga.gcb[e] += l * double(tp.y);
    }
}

// Transpose of the E update: lamH += sum of the hv/hu terms of the E edges that read each H.
//   Ex reads Hz(p)-Hz(p-y), Hy(p)-Hy(p-z); Ey reads Hx(p)-Hx(p-z), Hz(p)-Hz(p-x);
//   Ez reads Hy(p)-Hy(p-x), Hx(p)-Hx(p-y)
// so H at q gathers those terms at q and q+x, q+y, q+z. Then the step's gradient terms
// lamE^{n+1} (E^n, curl^n) are accumulated on the design edges (lamE is read-only here).
__global__ __launch_bounds__(256) void k_adj_e(Geo g, const float *__restrict__ la, const float *__restrict__ lamE,
                                               float *__restrict__ lamH, const float *__restrict__ lp, GradArgs ga) {
    const uint32_t k = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, i = blockIdx.z;
    if (k >= g.n2 || j >= g.n1) return;
    const uint64_t id = (uint64_t(i) * g.n1 + j) * g.n2 + k;
    float hv0, hu0, hv1, hu1, hv2, hu2, t, x0y, x0z, x1z, x1x, x2x, x2y;
    e_terms<0>(g, la, lp, i, j, k, hv0, hu0);
    e_terms<1>(g, la, lp, i, j, k, hv1, hu1);
    e_terms<2>(g, la, lp, i, j, k, hv2, hu2);
    e_terms<0>(g, la, lp, i, j + 1, k, x0y, t);  // hv of Ex at p+y
    e_terms<0>(g, la, lp, i, j, k + 1, t, x0z);  // hu of Ex at p+z
    e_terms<1>(g, la, lp, i, j, k + 1, x1z, t);  // hv of Ey at p+z
    e_terms<1>(g, la, lp, i + 1, j, k, t, x1x);  // hu of Ey at p+x
    e_terms<2>(g, la, lp, i + 1, j, k, x2x, t);  // hv of Ez at p+x
    e_terms<2>(g, la, lp, i, j + 1, k, t, x2y);  // hu of Ez at p+y
    lamH[id] += (hv1 - x1z) + (hu2 - x2y);
    lamH[g.nn + id] += (hv2 - x2x) + (hu0 - x0z);
    lamH[2 * g.nn + id] += (hv0 - x0y) + (hu1 - x1x);
    grad_terms(g, ga, lamE, i, j, k, id);
}

// psi adjoint of the node's own slots: psiBar_pre = be (psiBar + s) for an advanced slot, carried
// over unchanged where the forward step left the psi value alone.
template <int C>
__device__ __forceinline__ void psi_e_own(const Geo &g, float a, bool on, const float *__restrict__ lp,
                                          float *__restrict__ ln, uint32_t i, uint32_t j, uint32_t k) {
    constexpr int U = (C + 1) % 3, V = (C + 2) % 3;
    const uint32_t pu = coord<U>(i, j, k), pv = coord<V>(i, j, k);
    const int su = g.ax[U].slot_e[pu], sv = g.ax[V].slot_e[pv];
    if (su >= 0) {
        const uint64_t q = psi_at<U>(g, 0, C, i, j, k, uint32_t(su));
        ln[q] = on ? g.ax[U].be[pu] * (lp[q] + a) : lp[q];
    }
    if (sv >= 0) {
        const uint64_t q = psi_at<V>(g, 0, C, i, j, k, uint32_t(sv));
        ln[q] = on ? g.ax[V].be[pv] * (lp[q] - a) : lp[q];
    }
}

template <int C>
__device__ __forceinline__ void psi_h_own(const Geo &g, const float *__restrict__ lamH, const float *__restrict__ lp,
                                          float *__restrict__ ln, uint32_t i, uint32_t j, uint32_t k) {
    constexpr int U = (C + 1) % 3, V = (C + 2) % 3;
    const uint32_t pu = coord<U>(i, j, k), pv = coord<V>(i, j, k);
    const bool on = pu + 1 < dim<U>(g) && pv + 1 < dim<V>(g);
    const float b = -g.k_mu * lamH[C * g.nn + (uint64_t(i) * g.n1 + j) * g.n2 + k];
    const int su = g.ax[U].slot_h[pu], sv = g.ax[V].slot_h[pv];
    if (su >= 0) {
        const uint64_t q = psi_at<U>(g, 1, C, i, j, k, uint32_t(su));
        ln[q] = on ? g.ax[U].bh[pu] * (lp[q] + b) : lp[q];
    }
    if (sv >= 0) {
        const uint64_t q = psi_at<V>(g, 1, C, i, j, k, uint32_t(sv));
        ln[q] = on ? g.ax[V].bh[pv] * (lp[q] - b) : lp[q];
    }
}

// The observations whose dJ/ds enters lamE after this transpose (those of the step before).
struct InjectArgs {
    const uint64_t *__restrict__ obs;  // flat E index per observation
    const double *__restrict__ ds;     // dJ/ds of the step before, NULL at the first step
    uint32_t n_obs;
};

// Node (i,j,k)'s share of the H transpose, given the ev/eu terms at the node and at its -1
// neighbours: lamE = ca lamE + gathered terms (+ dJ/ds where observed), la = cb lamE, and the
// node's psi adjoints advanced into ln.
__device__ __forceinline__ void adj_h_node(const Geo &g, float *__restrict__ lamE, float *__restrict__ la,
                                           const float *__restrict__ lamH, const float *__restrict__ lp,
                                           float *__restrict__ ln, const InjectArgs &in, uint32_t i, uint32_t j,
                                           uint32_t k, const float t[6], const float y[6]) {
    // t = {ev0, eu0, ev1, eu1, ev2, eu2} at p; y = {ev0(p-y), eu0(p-z), ev1(p-z), eu1(p-x), ev2(p-x), eu2(p-y)}
    const uint64_t nn = g.nn, id = (uint64_t(i) * g.n1 + j) * g.n2 + k;
    float le[3] = {lamE[id], lamE[nn + id], lamE[2 * nn + id]};
    float a[3], cb[3];
    bool on[3];
    for (int c = 0; c < 3; c++) {
        const float2 cf = coef_at(g, c * nn + id);
        on[c] = cf.y != 0.f;
        cb[c] = cf.y;
        a[c] = cf.y * le[c];
        if (on[c]) le[c] *= cf.x;
    }
    le[0] += (y[2] - t[2]) + (y[5] - t[5]);  // Hy: ev on Ex (u = z); Hz: eu on Ex (v = y)
    le[1] += (y[4] - t[4]) + (y[1] - t[1]);  // Hz: ev on Ey (u = x); Hx: eu on Ey (v = z)
    le[2] += (y[0] - t[0]) + (y[3] - t[3]);  // Hx: ev on Ez (u = y); Hy: eu on Ez (v = x)
    if (in.ds)
        for (uint32_t q = 0; q < in.n_obs; q++) {
            const uint64_t f = in.obs[q];
            for (int c = 0; c < 3; c++)
                if (f == c * nn + id) le[c] += float(in.ds[q]);
        }
    for (int c = 0; c < 3; c++) {
        lamE[c * nn + id] = le[c];
        la[c * nn + id] = cb[c] * le[c];
    }
    if (g.pml) {
        psi_e_own<0>(g, a[0], on[0], lp, ln, i, j, k);
        psi_e_own<1>(g, a[1], on[1], lp, ln, i, j, k);
        psi_e_own<2>(g, a[2], on[2], lp, ln, i, j, k);
        psi_h_own<0>(g, lamH, lp, ln, i, j, k);
        psi_h_own<1>(g, lamH, lp, ln, i, j, k);
        psi_h_own<2>(g, lamH, lp, ln, i, j, k);
    }
}

// Transpose of the H update, and the rest of the E update's: lamE = ca lamE + the ev/eu terms of
// the H components that read each E (Hx reads Ez(p+y)-Ez(p), Ey(p+z)-Ey(p); Hy reads
// Ex(p+z)-Ex(p), Ez(p+x)-Ez(p); Hz reads Ey(p+x)-Ey(p), Ex(p+y)-Ex(p)), gathered at q and
// q-x, q-y, q-z; every psi adjoint advances from lp into ln; then dJ/ds of the step before
// enters lamE.
__global__ __launch_bounds__(256) void k_adj_h(Geo g, float *__restrict__ lamE, float *__restrict__ la,
                                               const float *__restrict__ lamH, const float *__restrict__ lp,
                                               float *__restrict__ ln, InjectArgs in) {
    const uint32_t k = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, i = blockIdx.z;
    if (k >= g.n2 || j >= g.n1) return;
    float t[6], y[6], u;
    h_terms<0>(g, lamH, lp, i, j, k, t[0], t[1]);
    h_terms<1>(g, lamH, lp, i, j, k, t[2], t[3]);
    h_terms<2>(g, lamH, lp, i, j, k, t[4], t[5]);
    h_terms<0>(g, lamH, lp, i, j - 1, k, y[0], u);  // ev of Hx at p-y
    h_terms<0>(g, lamH, lp, i, j, k - 1, u, y[1]);  // eu of Hx at p-z
    h_terms<1>(g, lamH, lp, i, j, k - 1, y[2], u);  // ev of Hy at p-z
    h_terms<1>(g, lamH, lp, i - 1, j, k, u, y[3]);  // eu of Hy at p-x
    h_terms<2>(g, lamH, lp, i - 1, j, k, y[4], u);  // ev of Hz at p-x
    h_terms<2>(g, lamH, lp, i, j - 1, k, u, y[5]);  // eu of Hz at p-y
    adj_h_node(g, lamE, la, lamH, lp, ln, in, i, j, k, t, y);
}

// ---- fused adjoint march: one pass per adjoint step ----------------------------------------
//
// The transposed step has the forward step's shape with the roles reversed: lamH gathers the
// hv/hu terms of lamE at the node and its +1 neighbours (as H gathers E), then lamE gathers the
// ev/eu terms of the new lamH at the node and its -1 neighbours (as E gathers H). So it marches
// like k_march: a block of TJ+2 rows x 64 lanes (rows 1..TJ and lanes 1..62 owned) walks a
// segment of i-planes; the +x terms are computed one plane ahead, the -x terms carried from the
// plane before, +/-z by lane shuffles, +/-y through LDS, one barrier per plane. lamE, lamH and
// the psi adjoints are ping-ponged (halo rows recompute from the old copies). The terms are the
// expressions of e_terms / h_terms / adj_h_node, so the result equals k_adj_e + k_adj_h bit for bit.
template <int TJ, int PML, bool WIDE, bool LDS_SETS>
__global__ __launch_bounds__(64 * (TJ + 2)) void k_adj_march(Geo g, const MarchBlock *__restrict__ blocks,
                                                             const float *__restrict__ Lo, float *__restrict__ Ln,
                                                             const float *__restrict__ Mo, float *__restrict__ Mn,
                                                             const float *__restrict__ po, float *__restrict__ pn,
                                                             uint32_t seg, GradArgs ga, InjectArgs in) {
    constexpr int ROWS = TJ + 2;
    constexpr bool PZ = PML >= 1, PXY = PML == 2;
    __shared__ float xv[2][ROWS + 2][2][64];  // hv of Ex, hu of Ez at a plane, per row (read by the row above)
    __shared__ float xw[2][ROWS + 2][2][64];  // ev of Hx, eu of Hz at a plane, per row (read by the row below)
    __shared__ float2 sets[LDS_SETS ? SETS_LDS : 1];
    __shared__ float xa[SEG_MAX + 2][PXY ? 14 : 4];
    const MarchBlock mb = blocks[blockIdx.x];
    if (mb.bx == ~0u) return;
    const int lane = threadIdx.x, w = __builtin_amdgcn_readfirstlane(threadIdx.y);
    const int k = int(mb.bx) * OWN_K + lane - 1;
    const int j = int(mb.by) * TJ + w - 1;
    const uint32_t i0 = mb.bz * seg, i1 = min(i0 + seg, g.n0);
    const bool kin = k >= 0 && k < int(g.n2), jin = j >= 0 && j < int(g.n1);
    const bool act = kin && jin;
    const bool own = act && lane >= 1 && lane <= OWN_K && w >= 1 && w <= TJ;
    const uint32_t uk = kin ? uint32_t(k) : 0u, uj = jin ? uint32_t(j) : 0u;
    const uint64_t nn = g.nn, plane = uint64_t(g.n1) * g.n2;
    const uint32_t off = uj * g.n2 + uk;
    const bool kl = uk + 1 < g.n2, jl = uj + 1 < g.n1;
    const float idz = g.ax[2].inv_d[uk], iddz = g.ax[2].inv_dd[uk];
    const float idy = g.ax[1].inv_d[uj], iddy = g.ax[1].inv_dd[uj];
    const uint32_t ob4 = off * 4u;
    const uint32_t ldA = act ? ob4 : OOB, stO = own ? ob4 : OOB;
    const uint32_t ldC = act ? off * (WIDE ? 4u : 2u) : OOB;
    const uint32_t pl4 = uint32_t(plane) * 4u, plc = uint32_t(plane) * (WIDE ? 4u : 2u);
    float bhz = 0, chz = 0, rdz = 0, bez = 0, cez = 0, rddz = 0, bhy = 0, chy = 0, rdy = 0, bey = 0, cey = 0, rddy = 0;
    uint32_t ozh = OOB, oze = OOB, oyh = OOB, oye = OOB;
    const uint32_t p2 = 2u * g.pml;
    const uint32_t zbytes = g.n1 * p2 * 4u, ybytes = p2 * g.n2 * 4u;
    if constexpr (PML) {
        const int shz = g.ax[2].slot_h[uk], sez = g.ax[2].slot_e[uk];
        const int shy = g.ax[1].slot_h[uj], sey = g.ax[1].slot_e[uj];
        bhz = g.ax[2].bh[uk], chz = g.ax[2].ch[uk], rdz = g.ax[2].raw_d[uk];
        bez = g.ax[2].be[uk], cez = g.ax[2].ce[uk], rddz = g.ax[2].raw_dd[uk];
        bhy = g.ax[1].bh[uj], chy = g.ax[1].ch[uj], rdy = g.ax[1].raw_d[uj];
        bey = g.ax[1].be[uj], cey = g.ax[1].ce[uj], rddy = g.ax[1].raw_dd[uj];
        if (act) {
            ozh = shz >= 0 ? (uj * p2 + uint32_t(shz)) * 4u : OOB, oze = sez >= 0 ? (uj * p2 + uint32_t(sez)) * 4u : OOB;
            oyh = shy >= 0 ? (uint32_t(shy) * g.n2 + uk) * 4u : OOB, oye = sey >= 0 ? (uint32_t(sey) * g.n2 + uk) * 4u : OOB;
        }
    }
    const uint32_t *cix = g.coef_index;
    const uint16_t *cix16 = reinterpret_cast<const uint16_t *>(g.coef_index);
    auto cidx = [&](uint64_t ub, uint32_t bytes) {
        if constexpr (WIDE) return ld(cix + ub, ldC, bytes);
        else return uint32_t(ld(cix16 + ub, ldC, bytes));
    };
    auto coef = [&](uint32_t c) {
        if constexpr (LDS_SETS) return sets[c];
        else return g.coef_sets[c];
    };
    if constexpr (LDS_SETS)
        for (uint32_t t = w * 64 + lane; t < g.n_sets; t += 64 * ROWS) sets[t] = g.coef_sets[t];
    const uint32_t istart = i0 > 0 ? i0 - 1 : 0;
    for (uint32_t t = w * 64 + lane; t + istart <= i1; t += 64 * ROWS) {
        const uint32_t q = min(istart + t, g.n0 - 1);
        xa[t][0] = g.ax[0].inv_d[q], xa[t][1] = g.ax[0].inv_dd[q];
        xa[t][2] = __int_as_float(g.ax[0].slot_h[q]), xa[t][3] = __int_as_float(g.ax[0].slot_e[q]);
        if constexpr (PXY) {
            xa[t][4] = g.ax[0].bh[q], xa[t][5] = g.ax[0].ch[q], xa[t][6] = g.ax[0].raw_d[q];
            xa[t][7] = g.ax[0].be[q], xa[t][8] = g.ax[0].ce[q], xa[t][9] = g.ax[0].raw_dd[q];
        }
    }
    if (w < 2)
        for (int q = 0; q < 2; q++) {
            xv[q][w * (ROWS + 1)][0][lane] = 0.f, xv[q][w * (ROWS + 1)][1][lane] = 0.f;
            xw[q][w * (ROWS + 1)][0][lane] = 0.f, xw[q][w * (ROWS + 1)][1][lane] = 0.f;
        }
    __syncthreads();  // sets and xa (slot numbers become buffer bases) are read from here on
#define PSI(e, c, a) g.psi_off[(e) * 9 + (c) * 3 + (a)]
    // Loads run one plane ahead of their use, so their latency overlaps a whole iteration: at
    // iteration i the raw values of plane i+2 (lamE, coefficient-set indices, E-psi adjoints) and
    // of plane i+1 (lamH, H-psi adjoints) are issued, while plane i+1's terms are formed from the
    // raw values issued the iteration before.
    struct RawE {
        float lx, ly, lz;
        uint32_t c0, c1, c2;
        float pxy, pxz, pyz, pyx, pzx, pzy;  // E-psi adjoints (component, axis)
    };
    struct RawH {
        float hx, hy, hz;
        float qxy, qxz, qyz, qyx, qzx, qzy;  // H-psi adjoints
    };
    // the hv/hu terms of plane p from lamE, (ca, cb) and the E-psi adjoints of that plane
    struct Ed {
        float lx, ly, lz;                          // lamE
        float2 f0, f1, f2;                         // (ca, cb)
        float pxy, pxz, pyz, pyx, pzx, pzy;        // E-psi adjoints (component, axis)
        float hv0, hu0, hv1, hu1, hv2, hu2;        // terms
    };
    auto load_e = [&](uint32_t p, RawE &d) {
        const bool in_ = p < g.n0;
        const float *q = Lo + (in_ ? uint64_t(p) * plane : 0);
        const uint32_t b = in_ ? pl4 : 0u;
        d.lx = ld(q, ldA, b), d.ly = ld(q + nn, ldA, b), d.lz = ld(q + 2 * nn, ldA, b);
        const uint64_t cb_ = in_ ? uint64_t(p) * plane : 0;
        const uint32_t cbytes = in_ ? plc : 0u;
        d.c0 = cidx(cb_, cbytes), d.c1 = cidx(nn + cb_, cbytes), d.c2 = cidx(2 * nn + cb_, cbytes);
        d.pxy = d.pxz = d.pyz = d.pyx = d.pzx = d.pzy = 0.f;
        if constexpr (PZ) {
            const uint64_t zrow = in_ ? uint64_t(p) * g.n1 * p2 : 0;
            const uint32_t zb = in_ ? zbytes : 0u;
            d.pxz = ld(po + PSI(0, 0, 2) + zrow, oze, zb), d.pyz = ld(po + PSI(0, 1, 2) + zrow, oze, zb);
        }
        if constexpr (PXY) {
            const float *xs = xa[min(p, i1) - istart];
            const int sex = __float_as_int(xs[3]);
            const uint64_t yrow = in_ ? uint64_t(p) * p2 * g.n2 : 0, xeb = uint64_t(max(sex, 0)) * plane;
            const uint32_t yb = in_ ? ybytes : 0u, xb = in_ ? pl4 : 0u, oxe = sex >= 0 ? ldA : OOB;
            d.pxy = ld(po + PSI(0, 0, 1) + yrow, oye, yb);
            d.pyx = ld(po + PSI(0, 1, 0) + xeb, oxe, xb);
            d.pzx = ld(po + PSI(0, 2, 0) + xeb, oxe, xb), d.pzy = ld(po + PSI(0, 2, 1) + yrow, oye, yb);
        }
    };
    auto load_h = [&](uint32_t i, RawH &d) {
        const uint64_t pb = uint64_t(i) * plane;
        d.hx = ld(Mo + pb, ldA, pl4), d.hy = ld(Mo + pb + nn, ldA, pl4), d.hz = ld(Mo + pb + 2 * nn, ldA, pl4);
        d.qxy = d.qxz = d.qyz = d.qyx = d.qzx = d.qzy = 0.f;
        if constexpr (PZ) {
            const uint64_t zrow = uint64_t(i) * g.n1 * p2;
            d.qxz = ld(po + PSI(1, 0, 2) + zrow, ozh, zbytes), d.qyz = ld(po + PSI(1, 1, 2) + zrow, ozh, zbytes);
        }
        if constexpr (PXY) {
            const float *xs = xa[i - istart];
            const int shx = __float_as_int(xs[2]);
            const uint64_t yrow = uint64_t(i) * p2 * g.n2, xhb = uint64_t(max(shx, 0)) * plane;
            const uint32_t oxh = shx >= 0 ? ldA : OOB;
            d.qxy = ld(po + PSI(1, 0, 1) + yrow, oyh, ybytes);
            d.qyx = ld(po + PSI(1, 1, 0) + xhb, oxh, pl4);
            d.qzx = ld(po + PSI(1, 2, 0) + xhb, oxh, pl4), d.qzy = ld(po + PSI(1, 2, 1) + yrow, oyh, ybytes);
        }
    };
    auto terms_e = [&](uint32_t p, const RawE &r, Ed &d) {
        d.lx = r.lx, d.ly = r.ly, d.lz = r.lz;
        d.f0 = coef(r.c0), d.f1 = coef(r.c1), d.f2 = coef(r.c2);
        d.pxy = r.pxy, d.pxz = r.pxz, d.pyz = r.pyz, d.pyx = r.pyx, d.pzx = r.pzx, d.pzy = r.pzy;
        const float *xs = xa[min(p, i1) - istart];
        const float iddx = xs[1];
        const float ax_ = d.f0.y != 0.f ? d.f0.y * d.lx : 0.f, ay_ = d.f1.y != 0.f ? d.f1.y * d.ly : 0.f,
                    az_ = d.f2.y != 0.f ? d.f2.y * d.lz : 0.f;
        d.hv0 = ax_ * iddy, d.hu0 = -ax_ * iddz;
        d.hv1 = ay_ * iddz, d.hu1 = -ay_ * iddx;
        d.hv2 = az_ * iddx, d.hu2 = -az_ * iddy;
        if constexpr (PZ) {
            d.hu0 += cez * rddz * (d.pxz - ax_);
            d.hv1 += cez * rddz * (d.pyz + ay_);
        }
        if constexpr (PXY) {
            const float cex = xs[8], rddx = xs[9];
            d.hv0 += cey * rddy * (d.pxy + ax_);
            d.hu1 += cex * rddx * (d.pyx - ay_);
            d.hv2 += cex * rddx * (d.pzx + az_);
            d.hu2 += cey * rddy * (d.pzy - az_);
        }
    };
    Ed cur, nxt;
    RawE re;
    RawH rh;
    load_e(istart, re);
    terms_e(istart, re, cur);
    load_e(istart + 1, re);  // istart < i1: the block has at least one plane
    load_h(istart, rh);
    xv[istart & 1][w + 1][0][lane] = cur.hv0, xv[istart & 1][w + 1][1][lane] = cur.hu2;
    __syncthreads();
    float evz_p = 0.f, euy_p = 0.f;  // ev of Hz, eu of Hy at the plane before (the -x terms)
    for (uint32_t i = istart; i < i1; i++) {
        const int b = i & 1;
        const uint64_t pb = uint64_t(i) * plane;
        const bool il = i + 1 < g.n0;
        const bool wplane = i >= i0;
        const uint32_t sto = wplane ? stO : OOB;
        const float *xs = xa[i - istart];
        const float idx_ = xs[0];
        const RawE ne = re;  // plane i+1, issued the iteration before
        const RawH h = rh;   // plane i
        terms_e(i + 1, ne, nxt);
        if (i + 1 < i1) {  // uniform
            load_e(i + 2, re);
            load_h(i + 1, rh);
        }
        const float hx0 = h.hx, hy0 = h.hy, hz0 = h.hz;
        const float qxy = h.qxy, qxz = h.qxz, qyz = h.qyz, qyx = h.qyx, qzx = h.qzx, qzy = h.qzy;
        uint64_t zrow = 0, yrow = 0, xhb = 0, xeb = 0;
        uint32_t oxh = OOB, oxe = OOB;
        float bhx = 0, chx = 0, rdx = 0, bex = 0;
        if constexpr (PZ) zrow = uint64_t(i) * g.n1 * p2;
        if constexpr (PXY) {
            const int shx = __float_as_int(xs[2]), sex = __float_as_int(xs[3]);
            bhx = xs[4], chx = xs[5], rdx = xs[6], bex = xs[7];
            yrow = uint64_t(i) * p2 * g.n2;
            xhb = uint64_t(max(shx, 0)) * plane, xeb = uint64_t(max(sex, 0)) * plane;
            oxh = shx >= 0 ? ldA : OOB, oxe = sex >= 0 ? ldA : OOB;
        }
        // lamH: + terms at the node, - terms at +z (next lane), +y (row below, LDS), +x (next plane)
        const float x1z = lane_next(cur.hv1), x0z = lane_next(cur.hu0);
        const float x0y = xv[b][w + 2][0][lane], x2y = xv[b][w + 2][1][lane];
        const float mx = hx0 + ((cur.hv1 - x1z) + (cur.hu2 - x2y));
        const float my = hy0 + ((cur.hv2 - nxt.hv2) + (cur.hu0 - x0z));
        const float mz = hz0 + ((cur.hv0 - x0y) + (cur.hu1 - nxt.hu1));
        {
            float *q = Mn + pb;
            st(q, sto, pl4, mx), st(q + nn, sto, pl4, my), st(q + 2 * nn, sto, pl4, mz);
        }
        // ev/eu terms of the new lamH: zero outside the grid (the halo's out-of-grid nodes hold
        // partial sums) and for the H components the forward step skips
        const float bx_ = act && jl && kl ? -g.k_mu * mx : 0.f, by_ = act && kl && il ? -g.k_mu * my : 0.f,
                    bz_ = act && il && jl ? -g.k_mu * mz : 0.f;
        float ev0 = bx_ * idy, eu0 = -bx_ * idz, ev1 = by_ * idz, eu1 = -by_ * idx_, ev2 = bz_ * idx_, eu2 = -bz_ * idy;
        if constexpr (PZ) {
            eu0 += chz * rdz * (qxz - bx_);
            ev1 += chz * rdz * (qyz + by_);
        }
        if constexpr (PXY) {
            ev0 += chy * rdy * (qxy + bx_);
            eu1 += chx * rdx * (qyx - by_);
            ev2 += chx * rdx * (qzx + bz_);
            eu2 += chy * rdy * (qzy - bz_);
        }
        xw[b][w + 1][0][lane] = ev0, xw[b][w + 1][1][lane] = eu2;
        xv[b ^ 1][w + 1][0][lane] = nxt.hv0, xv[b ^ 1][w + 1][1][lane] = nxt.hu2;
        __syncthreads();
        // lamE: ca lamE + terms at -z (previous lane), -y (row above, LDS), -x (plane before)
        const float y1z = lane_prev(ev1), y0z = lane_prev(eu0);
        const float y0y = xw[b][w][0][lane], y2y = xw[b][w][1][lane];
        float le0 = cur.f0.y != 0.f ? cur.lx * cur.f0.x : cur.lx;
        float le1 = cur.f1.y != 0.f ? cur.ly * cur.f1.x : cur.ly;
        float le2 = cur.f2.y != 0.f ? cur.lz * cur.f2.x : cur.lz;
        le0 += (y1z - ev1) + (y2y - eu2);
        le1 += (evz_p - ev2) + (y0z - eu0);
        le2 += (y0y - ev0) + (euy_p - eu1);
        if (own && wplane) {
            const uint64_t id = pb + off;
            if (in.ds)
                for (uint32_t q = 0; q < in.n_obs; q++) {
                    const uint64_t f = in.obs[q];
                    if (f == id) le0 += float(in.ds[q]);
                    if (f == nn + id) le1 += float(in.ds[q]);
                    if (f == 2 * nn + id) le2 += float(in.ds[q]);
                }
            // gradient terms of the design edges at this node, with the lamE the step started from
            const uint32_t bi = i - ga.lo[0], bj = uj - ga.lo[1], bk = uk - ga.lo[2];
            if (bi < ga.nb[0] && bj < ga.nb[1] && bk < ga.nb[2]) {
                const uint64_t bq = (uint64_t(bi) * ga.nb[1] + bj) * ga.nb[2] + bk,
                               nb = uint64_t(ga.nb[0]) * ga.nb[1] * ga.nb[2];
                const float lv[3] = {cur.lx, cur.ly, cur.lz};
                for (int c = 0; c < 3; c++) {
                    const int32_t e = ga.map[c * nb + bq];
                    if (e < 0) continue;
                    const double l = lv[c];
                    const float2 tp = ga.tape[e];
                    ga.gca[e] += l * double(tp.x);
                    ga.gcb[e] += l * double(tp.y);
                }
            }
        }
        {
            float *q = Ln + pb;
            st(q, sto, pl4, le0), st(q + nn, sto, pl4, le1), st(q + 2 * nn, sto, pl4, le2);
        }
        // psi adjoints of this plane into the other copy: advanced where the forward advanced them
        if constexpr (PZ) {
            const float ax_ = cur.f0.y * cur.lx, ay_ = cur.f1.y * cur.ly;
            const bool on0 = cur.f0.y != 0.f, on1 = cur.f1.y != 0.f;
            st(pn + PSI(0, 0, 2) + zrow, sto == OOB ? OOB : oze, zbytes, on0 ? bez * (cur.pxz - ax_) : cur.pxz);
            st(pn + PSI(0, 1, 2) + zrow, sto == OOB ? OOB : oze, zbytes, on1 ? bez * (cur.pyz + ay_) : cur.pyz);
            st(pn + PSI(1, 0, 2) + zrow, sto == OOB ? OOB : ozh, zbytes, jl && kl ? bhz * (qxz - bx_) : qxz);
            st(pn + PSI(1, 1, 2) + zrow, sto == OOB ? OOB : ozh, zbytes, kl && il ? bhz * (qyz + by_) : qyz);
        }
        if constexpr (PXY) {
            const float ax_ = cur.f0.y * cur.lx, ay_ = cur.f1.y * cur.ly, az_ = cur.f2.y * cur.lz;
            const bool on0 = cur.f0.y != 0.f, on1 = cur.f1.y != 0.f, on2 = cur.f2.y != 0.f;
            const uint32_t ye = sto == OOB ? OOB : oye, yh = sto == OOB ? OOB : oyh;
            const uint32_t xe_ = sto == OOB ? OOB : oxe, xh_ = sto == OOB ? OOB : oxh;
            st(pn + PSI(0, 0, 1) + yrow, ye, ybytes, on0 ? bey * (cur.pxy + ax_) : cur.pxy);
            st(pn + PSI(0, 1, 0) + xeb, xe_, pl4, on1 ? bex * (cur.pyx - ay_) : cur.pyx);
            st(pn + PSI(0, 2, 0) + xeb, xe_, pl4, on2 ? bex * (cur.pzx + az_) : cur.pzx);
            st(pn + PSI(0, 2, 1) + yrow, ye, ybytes, on2 ? bey * (cur.pzy - az_) : cur.pzy);
            st(pn + PSI(1, 0, 1) + yrow, yh, ybytes, jl && kl ? bhy * (qxy + bx_) : qxy);
            st(pn + PSI(1, 1, 0) + xhb, xh_, pl4, kl && il ? bhx * (qyx - by_) : qyx);
            st(pn + PSI(1, 2, 0) + xhb, xh_, pl4, il && jl ? bhx * (qzx + bz_) : qzx);
            st(pn + PSI(1, 2, 1) + yrow, yh, ybytes, il && jl ? bhy * (qzy - bz_) : qzy);
        }
        evz_p = ev2, euy_p = eu1;
        cur = nxt;
    }
#undef PSI
}

// ---- two adjoint steps per pass -----------------------------------------------------------
//
// k_adj_march2 runs adjoint steps n (A) and n-1 (B) in one pass over HBM, as march2 runs two
// forward steps: it reads lam^{n+1} from one ping-pong copy and writes lam^{n-1} into the other,
// and lam^n lives only in registers and LDS. A block is TJ+4 rows x 64 lanes (j0-2..j0+TJ+1,
// k0-2..k0+61) owning the inner TJ x 60; step A runs on the whole tile and step B on the tile
// less one ring, so B reads exactly the lam^n the owners of its neighbours computed. Along i
// the block marches with a one-plane lag: iteration t runs step A at plane t and step B at plane
// t-1. Every term is the expression k_adj_march evaluates, so two launches of k_adj_march and
// one of k_adj_march2 give the same bits; step A's dJ/ds enters every copy of its observed
// edges (owner or halo), and each design edge's owner adds A's gradient terms, then B's, in the
// order two one-step launches add them. The block classes read one plane past the segment's end
// (march_classes2 with reach 1): step A's lamH at the last plane needs lamE one plane on.
struct AdjE {  // the inputs of one plane's hv/hu terms
    float lx, ly, lz;
    float2 f0, f1, f2;
    float pxy, pxz, pyz, pyx, pzx, pzy;  // E-psi adjoints (component, axis)
};

template <int TJ, int PML, bool WIDE, bool LDS_SETS>
__global__ __launch_bounds__(64 * (TJ + 4)) void k_adj_march2(Geo g, const MarchBlock *__restrict__ blocks,
                                                              const float *__restrict__ Lo, float *__restrict__ Ln,
                                                              const float *__restrict__ Mo, float *__restrict__ Mn,
                                                              const float *__restrict__ po, float *__restrict__ pn,
                                                              uint32_t seg, GradArgs ga, const float2 *__restrict__ tapeB,
                                                              InjectArgs in, const double *__restrict__ dsB) {
    constexpr int ROWS = TJ + 4;
    constexpr bool PZ = PML >= 1, PXY = PML == 2;
    // [step A, B][plane parity][row + 1][field][lane]; rows -1 and ROWS are zero pads
    __shared__ float xv[2][2][ROWS + 2][2][64];  // hv of Ex, hu of Ez of a row (read by the row above)
    __shared__ float xw[2][2][ROWS + 2][2][64];  // ev of Hx, eu of Hz of a row (read by the row below)
    __shared__ float2 sets[LDS_SETS ? SETS_LDS : 1];
    __shared__ float xa[SEG_MAX + 4][PXY ? 10 : 2];
    const MarchBlock mb = blocks[blockIdx.x];
    if (mb.bx == ~0u) return;
    const int lane = threadIdx.x, w = __builtin_amdgcn_readfirstlane(threadIdx.y);
    const int k = int(mb.bx) * OWN_K2 + lane - 2;
    const int j = int(mb.by) * TJ + w - 2;
    const uint32_t i0 = mb.bz * seg, i1 = min(i0 + seg, g.n0);
    const bool kin = k >= 0 && k < int(g.n2), jin = j >= 0 && j < int(g.n1);
    const bool act = kin && jin;
    const bool own = act && lane >= 2 && lane < 2 + OWN_K2 && w >= 2 && w < 2 + TJ;
    const uint32_t uk = kin ? uint32_t(k) : 0u, uj = jin ? uint32_t(j) : 0u;
    const uint64_t nn = g.nn, plane = uint64_t(g.n1) * g.n2;
    const uint32_t off = uj * g.n2 + uk;
    const bool kl = uk + 1 < g.n2, jl = uj + 1 < g.n1;
    const float idz = g.ax[2].inv_d[uk], iddz = g.ax[2].inv_dd[uk];
    const float idy = g.ax[1].inv_d[uj], iddy = g.ax[1].inv_dd[uj];
    const uint32_t ob4 = off * 4u;
    const uint32_t ldA = act ? ob4 : OOB, stO = own ? ob4 : OOB;
    const uint32_t ldC = act ? off * (WIDE ? 4u : 2u) : OOB;
    const uint32_t pl4 = uint32_t(plane) * 4u, plc = uint32_t(plane) * (WIDE ? 4u : 2u);
    float bhz = 0, chz = 0, rdz = 0, bez = 0, cez = 0, rddz = 0, bhy = 0, chy = 0, rdy = 0, bey = 0, cey = 0, rddy = 0;
    uint32_t ozh = OOB, oze = OOB, oyh = OOB, oye = OOB;
    const uint32_t p2 = 2u * g.pml;
    const uint32_t zbytes = g.n1 * p2 * 4u, ybytes = p2 * g.n2 * 4u;
    if constexpr (PML) {
        const int shz = g.ax[2].slot_h[uk], sez = g.ax[2].slot_e[uk];
        const int shy = g.ax[1].slot_h[uj], sey = g.ax[1].slot_e[uj];
        bhz = g.ax[2].bh[uk], chz = g.ax[2].ch[uk], rdz = g.ax[2].raw_d[uk];
        bez = g.ax[2].be[uk], cez = g.ax[2].ce[uk], rddz = g.ax[2].raw_dd[uk];
        bhy = g.ax[1].bh[uj], chy = g.ax[1].ch[uj], rdy = g.ax[1].raw_d[uj];
        bey = g.ax[1].be[uj], cey = g.ax[1].ce[uj], rddy = g.ax[1].raw_dd[uj];
        if (act) {
            ozh = shz >= 0 ? (uj * p2 + uint32_t(shz)) * 4u : OOB, oze = sez >= 0 ? (uj * p2 + uint32_t(sez)) * 4u : OOB;
            oyh = shy >= 0 ? (uint32_t(shy) * g.n2 + uk) * 4u : OOB, oye = sey >= 0 ? (uint32_t(sey) * g.n2 + uk) * 4u : OOB;
        }
    }
    const uint32_t *cix = g.coef_index;
    const uint16_t *cix16 = reinterpret_cast<const uint16_t *>(g.coef_index);
    auto cidx = [&](uint64_t ub, uint32_t bytes) {
        if constexpr (WIDE) return ld(cix + ub, ldC, bytes);
        else return uint32_t(ld(cix16 + ub, ldC, bytes));
    };
    auto coef = [&](uint32_t c) {
        if constexpr (LDS_SETS) return sets[c];
        else return g.coef_sets[c];
    };
    if constexpr (LDS_SETS)
        for (uint32_t t = w * 64 + lane; t < g.n_sets; t += 64 * ROWS) sets[t] = g.coef_sets[t];
    // step A runs on planes tstart..i1 (its lamE on i0-1..i1 is what step B reads), step B on
    // tstart..i1-1 and stores i0..i1-1; x-axis values of planes tstart..i1+1
    const uint32_t tstart = i0 >= 2 ? i0 - 2 : 0;
    for (uint32_t t = w * 64 + lane; t + tstart <= i1 + 1; t += 64 * ROWS) {
        const uint32_t q = min(tstart + t, g.n0 - 1);
        xa[t][0] = g.ax[0].inv_d[q], xa[t][1] = g.ax[0].inv_dd[q];
        if constexpr (PXY) {
            xa[t][2] = g.ax[0].bh[q], xa[t][3] = g.ax[0].ch[q], xa[t][4] = g.ax[0].raw_d[q];
            xa[t][5] = g.ax[0].be[q], xa[t][6] = g.ax[0].ce[q], xa[t][7] = g.ax[0].raw_dd[q];
            xa[t][8] = __int_as_float(g.ax[0].slot_h[q]), xa[t][9] = __int_as_float(g.ax[0].slot_e[q]);
        }
    }
    if (w < 2)
        for (int a = 0; a < 2; a++)
            for (int q = 0; q < 2; q++) {
                xv[a][q][w * (ROWS + 1)][0][lane] = 0.f, xv[a][q][w * (ROWS + 1)][1][lane] = 0.f;
                xw[a][q][w * (ROWS + 1)][0][lane] = 0.f, xw[a][q][w * (ROWS + 1)][1][lane] = 0.f;
            }
    __syncthreads();  // sets and xa (slot numbers become buffer bases) are read from here on
#define PSI(e, c, a) g.psi_off[(e) * 9 + (c) * 3 + (a)]
    // raw loads, issued one plane ahead of their use as in k_adj_march
    struct RawE {
        float lx, ly, lz;
        uint32_t c0, c1, c2;
        float pxy, pxz, pyz, pyx, pzx, pzy;
    };
    struct RawH {
        float hx, hy, hz;
        float qxy, qxz, qyz, qyx, qzx, qzy;  // H-psi adjoints
    };
    struct Ed {  // AdjE plus the terms it gives
        AdjE a;
        float hv0, hu0, hv1, hu1, hv2, hu2;
    };
    auto load_e = [&](uint32_t p, RawE &d) {
        const bool in_ = p < g.n0;
        const float *q = Lo + (in_ ? uint64_t(p) * plane : 0);
        const uint32_t b = in_ ? pl4 : 0u;
        d.lx = ld(q, ldA, b), d.ly = ld(q + nn, ldA, b), d.lz = ld(q + 2 * nn, ldA, b);
        const uint64_t cb_ = in_ ? uint64_t(p) * plane : 0;
        const uint32_t cbytes = in_ ? plc : 0u;
        d.c0 = cidx(cb_, cbytes), d.c1 = cidx(nn + cb_, cbytes), d.c2 = cidx(2 * nn + cb_, cbytes);
        d.pxy = d.pxz = d.pyz = d.pyx = d.pzx = d.pzy = 0.f;
        if constexpr (PZ) {
            const uint64_t zrow = in_ ? uint64_t(p) * g.n1 * p2 : 0;
            const uint32_t zb = in_ ? zbytes : 0u;
            d.pxz = ld(po + PSI(0, 0, 2) + zrow, oze, zb), d.pyz = ld(po + PSI(0, 1, 2) + zrow, oze, zb);
        }
        if constexpr (PXY) {
            const float *xs = xa[min(p, i1 + 1) - tstart];
            const int sex = __float_as_int(xs[9]);
            const uint64_t yrow = in_ ? uint64_t(p) * p2 * g.n2 : 0, xeb = uint64_t(max(sex, 0)) * plane;
            const uint32_t yb = in_ ? ybytes : 0u, xb = in_ ? pl4 : 0u, oxe = sex >= 0 ? ldA : OOB;
            d.pxy = ld(po + PSI(0, 0, 1) + yrow, oye, yb);
            d.pyx = ld(po + PSI(0, 1, 0) + xeb, oxe, xb);
            d.pzx = ld(po + PSI(0, 2, 0) + xeb, oxe, xb), d.pzy = ld(po + PSI(0, 2, 1) + yrow, oye, yb);
        }
    };
    // plane i of the adjoint H; a plane past the grid (the last segment's i1 = n0) reads zeros
    // through empty resources
    auto load_h = [&](uint32_t i, RawH &d) {
        const bool in_ = i < g.n0;
        const uint64_t pb = in_ ? uint64_t(i) * plane : 0;
        const uint32_t hb = in_ ? pl4 : 0u;
        d.hx = ld(Mo + pb, ldA, hb), d.hy = ld(Mo + pb + nn, ldA, hb), d.hz = ld(Mo + pb + 2 * nn, ldA, hb);
        d.qxy = d.qxz = d.qyz = d.qyx = d.qzx = d.qzy = 0.f;
        if constexpr (PZ) {
            const uint64_t zrow = in_ ? uint64_t(i) * g.n1 * p2 : 0;
            const uint32_t zb = in_ ? zbytes : 0u;
            d.qxz = ld(po + PSI(1, 0, 2) + zrow, ozh, zb), d.qyz = ld(po + PSI(1, 1, 2) + zrow, ozh, zb);
        }
        if constexpr (PXY) {
            const float *xs = xa[i - tstart];
            const int shx = __float_as_int(xs[8]);
            const uint64_t yrow = in_ ? uint64_t(i) * p2 * g.n2 : 0, xhb = uint64_t(max(shx, 0)) * plane;
            const uint32_t oxh = shx >= 0 ? ldA : OOB, yb = in_ ? ybytes : 0u, xb = in_ ? pl4 : 0u;
            d.qxy = ld(po + PSI(1, 0, 1) + yrow, oyh, yb);
            d.qyx = ld(po + PSI(1, 1, 0) + xhb, oxh, xb);
            d.qzx = ld(po + PSI(1, 2, 0) + xhb, oxh, xb), d.qzy = ld(po + PSI(1, 2, 1) + yrow, oyh, yb);
        }
    };
    auto from_raw = [&](const RawE &r) {
        return AdjE{r.lx, r.ly, r.lz, coef(r.c0), coef(r.c1), coef(r.c2), r.pxy, r.pxz, r.pyz, r.pyx, r.pzx, r.pzy};
    };
    // the hv/hu terms of plane p (k_adj_march's terms_e)
    auto terms = [&](uint32_t p, const AdjE &a, Ed &d) {
        d.a = a;
        const float *xs = xa[min(p, i1 + 1) - tstart];
        const float iddx = xs[1];
        const float ax_ = a.f0.y != 0.f ? a.f0.y * a.lx : 0.f, ay_ = a.f1.y != 0.f ? a.f1.y * a.ly : 0.f,
                    az_ = a.f2.y != 0.f ? a.f2.y * a.lz : 0.f;
        d.hv0 = ax_ * iddy, d.hu0 = -ax_ * iddz;
        d.hv1 = ay_ * iddz, d.hu1 = -ay_ * iddx;
        d.hv2 = az_ * iddx, d.hu2 = -az_ * iddy;
        if constexpr (PZ) {
            d.hu0 += cez * rddz * (a.pxz - ax_);
            d.hv1 += cez * rddz * (a.pyz + ay_);
        }
        if constexpr (PXY) {
            const float cex = xs[6], rddx = xs[7];
            d.hv0 += cey * rddy * (a.pxy + ax_);
            d.hu1 += cex * rddx * (a.pyx - ay_);
            d.hv2 += cex * rddx * (a.pzx + az_);
            d.hu2 += cey * rddy * (a.pzy - az_);
        }
    };
    // One transposed step at plane p on the terms of p (cur) and p+1 (nxt), the adjoint H of p
    // (h: lamH and H-psi adjoints) and the -x carries; `pa` the step's LDS slot. Phase 1 forms
    // lamH and its ev/eu terms; after the caller's barrier, phase 2 forms lamE and the psi
    // adjoints of the step's output.
    struct HOut {
        float mx, my, mz;            // new lamH
        float bx, by, bz;            // -k_mu lamH where the forward updates it, else 0
        float ev0, eu0, ev1, eu1, ev2, eu2;
        float qxy, qxz, qyz, qyx, qzx, qzy;  // new H-psi adjoints
    };
    auto phase1 = [&](int pa, int b, uint32_t p, const Ed &cur, const Ed &nxt, const RawH &h, HOut &o) {
        const float *xs = xa[p - tstart];
        const float idx_ = xs[0];
        const bool il = p + 1 < g.n0;
        const float x1z = lane_next(cur.hv1), x0z = lane_next(cur.hu0);
        const float x0y = xv[pa][b][w + 2][0][lane], x2y = xv[pa][b][w + 2][1][lane];
        o.mx = h.hx + ((cur.hv1 - x1z) + (cur.hu2 - x2y));
        o.my = h.hy + ((cur.hv2 - nxt.hv2) + (cur.hu0 - x0z));
        o.mz = h.hz + ((cur.hv0 - x0y) + (cur.hu1 - nxt.hu1));
        o.bx = act && jl && kl ? -g.k_mu * o.mx : 0.f, o.by = act && kl && il ? -g.k_mu * o.my : 0.f,
        o.bz = act && il && jl ? -g.k_mu * o.mz : 0.f;
        o.ev0 = o.bx * idy, o.eu0 = -o.bx * idz, o.ev1 = o.by * idz, o.eu1 = -o.by * idx_, o.ev2 = o.bz * idx_,
        o.eu2 = -o.bz * idy;
        float bhx = 0, chx = 0, rdx = 0;
        if constexpr (PXY) bhx = xs[2], chx = xs[3], rdx = xs[4];
        if constexpr (PZ) {
            o.eu0 += chz * rdz * (h.qxz - o.bx);
            o.ev1 += chz * rdz * (h.qyz + o.by);
        }
        if constexpr (PXY) {
            o.ev0 += chy * rdy * (h.qxy + o.bx);
            o.eu1 += chx * rdx * (h.qyx - o.by);
            o.ev2 += chx * rdx * (h.qzx + o.bz);
            o.eu2 += chy * rdy * (h.qzy - o.bz);
        }
        o.qxy = h.qxy, o.qxz = h.qxz, o.qyz = h.qyz, o.qyx = h.qyx, o.qzx = h.qzx, o.qzy = h.qzy;
        if constexpr (PZ) {
            o.qxz = jl && kl ? bhz * (h.qxz - o.bx) : h.qxz;
            o.qyz = kl && il ? bhz * (h.qyz + o.by) : h.qyz;
        }
        if constexpr (PXY) {
            o.qxy = jl && kl ? bhy * (h.qxy + o.bx) : h.qxy;
            o.qyx = kl && il ? bhx * (h.qyx - o.by) : h.qyx;
            o.qzx = il && jl ? bhx * (h.qzx + o.bz) : h.qzx;
            o.qzy = il && jl ? bhy * (h.qzy - o.bz) : h.qzy;
        }
        xw[pa][b][w + 1][0][lane] = o.ev0, xw[pa][b][w + 1][1][lane] = o.eu2;
    };
    // phase 2: lamE of plane p (before dJ/ds) into out.lx..lz, the E-psi adjoints into out.p*
    auto phase2 = [&](int pa, int b, uint32_t p, const Ed &cur, const HOut &o, float evz_p, float euy_p, AdjE &out) {
        const float y1z = lane_prev(o.ev1), y0z = lane_prev(o.eu0);
        const float y0y = xw[pa][b][w][0][lane], y2y = xw[pa][b][w][1][lane];
        const AdjE &a = cur.a;
        float le0 = a.f0.y != 0.f ? a.lx * a.f0.x : a.lx;
        float le1 = a.f1.y != 0.f ? a.ly * a.f1.x : a.ly;
        float le2 = a.f2.y != 0.f ? a.lz * a.f2.x : a.lz;
        le0 += (y1z - o.ev1) + (y2y - o.eu2);
        le1 += (evz_p - o.ev2) + (y0z - o.eu0);
        le2 += (y0y - o.ev0) + (euy_p - o.eu1);
        out.lx = le0, out.ly = le1, out.lz = le2;
        out.f0 = a.f0, out.f1 = a.f1, out.f2 = a.f2;
        out.pxy = a.pxy, out.pxz = a.pxz, out.pyz = a.pyz, out.pyx = a.pyx, out.pzx = a.pzx, out.pzy = a.pzy;
        if constexpr (PZ) {
            const float ax_ = a.f0.y * a.lx, ay_ = a.f1.y * a.ly;
            out.pxz = a.f0.y != 0.f ? bez * (a.pxz - ax_) : a.pxz;
            out.pyz = a.f1.y != 0.f ? bez * (a.pyz + ay_) : a.pyz;
        }
        if constexpr (PXY) {
            const float *xs = xa[p - tstart];
            const float bex = xs[5];
            const float ax_ = a.f0.y * a.lx, ay_ = a.f1.y * a.ly, az_ = a.f2.y * a.lz;
            out.pxy = a.f0.y != 0.f ? bey * (a.pxy + ax_) : a.pxy;
            out.pyx = a.f1.y != 0.f ? bex * (a.pyx - ay_) : a.pyx;
            out.pzx = a.f2.y != 0.f ? bex * (a.pzx + az_) : a.pzx;
            out.pzy = a.f2.y != 0.f ? bey * (a.pzy - az_) : a.pzy;
        }
    };
    // dJ/ds into lamE at node id of plane p
    auto inject = [&](const double *ds, uint64_t id, AdjE &out) {
        if (!ds) return;
        for (uint32_t q = 0; q < in.n_obs; q++) {
            const uint64_t f = in.obs[q];
            if (f == id) out.lx += float(ds[q]);
            if (f == nn + id) out.ly += float(ds[q]);
            if (f == 2 * nn + id) out.lz += float(ds[q]);
        }
    };
    // the design edges' gradient terms at node (p, uj, uk), with the lamE the step started from
    auto grad = [&](uint32_t p, const AdjE &a, const float2 *__restrict__ tape) {
        const uint32_t bi = p - ga.lo[0], bj = uj - ga.lo[1], bk = uk - ga.lo[2];
        if (bi < ga.nb[0] && bj < ga.nb[1] && bk < ga.nb[2]) {
            const uint64_t bq = (uint64_t(bi) * ga.nb[1] + bj) * ga.nb[2] + bk,
                           nb = uint64_t(ga.nb[0]) * ga.nb[1] * ga.nb[2];
            const float lv[3] = {a.lx, a.ly, a.lz};
            for (int c = 0; c < 3; c++) {
                const int32_t e = ga.map[c * nb + bq];
                if (e < 0) continue;
                const double l = lv[c];
                const float2 tp = tape[e];
                ga.gca[e] += l * double(tp.x);
                ga.gcb[e] += l * double(tp.y);
            }
        }
    };
    Ed curA, nxtA, curB, newB;
    RawE re;
    RawH rh;
    load_e(tstart, re);
    terms(tstart, from_raw(re), curA);
    load_e(tstart + 1, re);  // tstart < i1: the block has at least one plane
    load_h(tstart, rh);
    xv[0][tstart & 1][w + 1][0][lane] = curA.hv0, xv[0][tstart & 1][w + 1][1][lane] = curA.hu2;
    __syncthreads();
    float evz_a = 0.f, euy_a = 0.f, evz_b = 0.f, euy_b = 0.f;  // the -x terms (previous plane) of A and B
    RawH hB{};  // step A's lamH and H-psi adjoints of the previous plane: step B's input there
    curB.a = AdjE{};
    curB.hv0 = curB.hu0 = curB.hv1 = curB.hu1 = curB.hv2 = curB.hu2 = 0.f;
    for (uint32_t t = tstart; t <= i1; t++) {
        const int b = t & 1;
        const RawE ne = re;  // plane t+1, issued the iteration before
        const RawH h = rh;   // plane t
        terms(t + 1, from_raw(ne), nxtA);
        if (t < i1) {  // uniform
            load_e(t + 2, re);
            load_h(t + 1, rh);
        }
        // ---- step A at plane t, phase 1
        HOut oa;
        phase1(0, b, t, curA, nxtA, h, oa);
        xv[0][b ^ 1][w + 1][0][lane] = nxtA.hv0, xv[0][b ^ 1][w + 1][1][lane] = nxtA.hu2;
        __syncthreads();
        // ---- step A at plane t, phase 2: lam^n of plane t, every copy
        AdjE la;
        phase2(0, b, t, curA, oa, evz_a, euy_a, la);
        if (t >= g.n0) la = AdjE{};  // no plane there: its terms are the zeros k_adj_march loads
        if (act && t < g.n0) inject(in.ds, uint64_t(t) * plane + off, la);
        if (own && t >= i0 && t < i1) grad(t, curA.a, ga.tape);
        evz_a = oa.ev2, euy_a = oa.eu1;
        terms(t, la, newB);
        // ---- step B at plane s = t-1, phase 1
        const uint32_t s = t - 1;
        const int sb = s & 1;
        const bool two = t > tstart;  // uniform
        HOut ob;
        if (two) {
            phase1(1, sb, s, curB, newB, hB, ob);
            const bool wr = own && s >= i0;
            const uint32_t sto = wr ? stO : OOB;
            float *q = Mn + uint64_t(s) * plane;
            st(q, sto, pl4, ob.mx), st(q + nn, sto, pl4, ob.my), st(q + 2 * nn, sto, pl4, ob.mz);
            if constexpr (PZ) {
                const uint64_t zrow = uint64_t(s) * g.n1 * p2;
                const uint32_t zh = sto == OOB ? OOB : ozh;
                st(pn + PSI(1, 0, 2) + zrow, zh, zbytes, ob.qxz), st(pn + PSI(1, 1, 2) + zrow, zh, zbytes, ob.qyz);
            }
            if constexpr (PXY) {
                const float *xs = xa[s - tstart];
                const int shx = __float_as_int(xs[8]);
                const uint64_t yrow = uint64_t(s) * p2 * g.n2, xhb = uint64_t(max(shx, 0)) * plane;
                const uint32_t yh = sto == OOB ? OOB : oyh, xh_ = sto == OOB || shx < 0 ? OOB : ldA;
                st(pn + PSI(1, 0, 1) + yrow, yh, ybytes, ob.qxy);
                st(pn + PSI(1, 1, 0) + xhb, xh_, pl4, ob.qyx);
                st(pn + PSI(1, 2, 0) + xhb, xh_, pl4, ob.qzx);
                st(pn + PSI(1, 2, 1) + yrow, yh, ybytes, ob.qzy);
            }
        }
        xv[1][b][w + 1][0][lane] = newB.hv0, xv[1][b][w + 1][1][lane] = newB.hu2;
        __syncthreads();
        // ---- step B at plane s, phase 2: lam^{n-1} of plane s, stored by its owner
        if (two) {
            AdjE lb;
            phase2(1, sb, s, curB, ob, evz_b, euy_b, lb);
            const bool wr = own && s >= i0;
            if (wr) {
                inject(dsB, uint64_t(s) * plane + off, lb);
                grad(s, curB.a, tapeB);
            }
            const uint32_t sto = wr ? stO : OOB;
            float *q = Ln + uint64_t(s) * plane;
            st(q, sto, pl4, lb.lx), st(q + nn, sto, pl4, lb.ly), st(q + 2 * nn, sto, pl4, lb.lz);
            if constexpr (PZ) {
                const uint64_t zrow = uint64_t(s) * g.n1 * p2;
                const uint32_t ze = sto == OOB ? OOB : oze;
                st(pn + PSI(0, 0, 2) + zrow, ze, zbytes, lb.pxz), st(pn + PSI(0, 1, 2) + zrow, ze, zbytes, lb.pyz);
            }
            if constexpr (PXY) {
                const float *xs = xa[s - tstart];
                const int sex = __float_as_int(xs[9]);
                const uint64_t yrow = uint64_t(s) * p2 * g.n2, xeb = uint64_t(max(sex, 0)) * plane;
                const uint32_t ye = sto == OOB ? OOB : oye, xe_ = sto == OOB || sex < 0 ? OOB : ldA;
                st(pn + PSI(0, 0, 1) + yrow, ye, ybytes, lb.pxy);
                st(pn + PSI(0, 1, 0) + xeb, xe_, pl4, lb.pyx);
                st(pn + PSI(0, 2, 0) + xeb, xe_, pl4, lb.pzx);
                st(pn + PSI(0, 2, 1) + yrow, ye, ybytes, lb.pzy);
            }
            evz_b = ob.ev2, euy_b = ob.eu1;
        }
        // rotate: plane t's step-A results become step B's input at the next iteration
        hB.hx = oa.mx, hB.hy = oa.my, hB.hz = oa.mz;
        hB.qxy = oa.qxy, hB.qxz = oa.qxz, hB.qyz = oa.qyz, hB.qyx = oa.qyx, hB.qzx = oa.qzx, hB.qzy = oa.qzy;
        curA = nxtA;
        curB = newB;
    }
#undef PSI
}

// dJ/ds_n enters lamE at the observed edges (s_n is E after step n)
__global__ void k_adj_inject(Geo g, float *__restrict__ lamE, float *__restrict__ la,
                             const uint64_t *__restrict__ obs, uint32_t n_obs, const double *__restrict__ ds) {
    if (threadIdx.x) return;
    for (uint32_t q = 0; q < n_obs; q++) {  // one thread: observations may share an edge
        const uint64_t f = obs[q];
        lamE[f] += float(ds[q]);
        if (la) la[f] = coef_at(g, f).y * lamE[f];
    }
}

// The forward step's E^n and curl^n = (t1 - t2) on the design edges (no CPML there), from the
// previous E copy and the new H copy of the ping-pong pair.
__global__ __launch_bounds__(256) void k_adj_tape(Geo g, const uint64_t *__restrict__ edges, uint32_t ne,
                                                  const float *__restrict__ Eo, const float *__restrict__ Hn,
                                                  float2 *__restrict__ tape) {
    const uint32_t e = blockIdx.x * 256 + threadIdx.x;
    if (e >= ne) return;
    const uint64_t f = edges[e], nn = g.nn;
    const uint32_t c = uint32_t(f / nn);
    const uint64_t id = f - c * nn;
    const uint32_t k = uint32_t(id % g.n2), j = uint32_t((id / g.n2) % g.n1), i = uint32_t(id / (uint64_t(g.n1) * g.n2));
    const uint64_t sx = uint64_t(g.n1) * g.n2, sy = g.n2;
    float curl;
    if (c == 0) {  // u = y, v = z: hv = Hz(p)-Hz(p-y), hu = Hy(p)-Hy(p-z)
        curl = (Hn[2 * nn + id] - Hn[2 * nn + id - sy]) * g.ax[1].inv_dd[j] - (Hn[nn + id] - Hn[nn + id - 1]) * g.ax[2].inv_dd[k];
    } else if (c == 1) {  // u = z, v = x: hv = Hx(p)-Hx(p-z), hu = Hz(p)-Hz(p-x)
        curl = (Hn[id] - Hn[id - 1]) * g.ax[2].inv_dd[k] - (Hn[2 * nn + id] - Hn[2 * nn + id - sx]) * g.ax[0].inv_dd[i];
    } else {  // u = x, v = y: hv = Hy(p)-Hy(p-x), hu = Hx(p)-Hx(p-y)
        curl = (Hn[nn + id] - Hn[nn + id - sx]) * g.ax[0].inv_dd[i] - (Hn[id] - Hn[id - sy]) * g.ax[1].inv_dd[j];
    }
    tape[e] = make_float2(Eo[f], curl);
}

// The tape of both steps of a two-step launch: step n's E^n is in the copy the launch read (Eo)
// and its curl^n comes from H^{n+1/2} recomputed from that copy; step n+1's E^{n+1} is
// recomputed from it the same way, and its curl^{n+1} comes from the new H copy (Hn).
__global__ __launch_bounds__(256) void k_adj_tape2(Geo g, FixArgs f, const uint64_t *__restrict__ edges, uint32_t ne,
                                                   const float *__restrict__ Eo, const float *__restrict__ Ho,
                                                   const float *__restrict__ Po, const float *__restrict__ Hn,
                                                   uint32_t s_off, float2 *__restrict__ tape0, float2 *__restrict__ tape1) {
    const uint32_t e = blockIdx.x * 256 + threadIdx.x;
    if (e >= ne) return;
    const uint64_t fl = edges[e], nn = g.nn;
    const uint32_t c = uint32_t(fl / nn);
    const uint64_t id = fl - c * nn;
    const uint32_t k = uint32_t(id % g.n2), j = uint32_t((id / g.n2) % g.n1), i = uint32_t(id / (uint64_t(g.n1) * g.n2));
    const uint32_t n = f.step[0] + s_off;
    const float pl = n < f.n_pulse ? f.pulse[n] : pulse_at(f, (float(n) + 0.5f) * f.dt);
    float curl0;
    if (c == 0) {  // u = y, v = z: hv = Hz(p)-Hz(p-y), hu = Hy(p)-Hy(p-z)
        curl0 = (h_mid<2>(g, Eo, Ho, Po, i, j, k) - h_mid<2>(g, Eo, Ho, Po, i, j - 1, k)) * g.ax[1].inv_dd[j] -
                (h_mid<1>(g, Eo, Ho, Po, i, j, k) - h_mid<1>(g, Eo, Ho, Po, i, j, k - 1)) * g.ax[2].inv_dd[k];
    } else if (c == 1) {  // u = z, v = x: hv = Hx(p)-Hx(p-z), hu = Hz(p)-Hz(p-x)
        curl0 = (h_mid<0>(g, Eo, Ho, Po, i, j, k) - h_mid<0>(g, Eo, Ho, Po, i, j, k - 1)) * g.ax[2].inv_dd[k] -
                (h_mid<2>(g, Eo, Ho, Po, i, j, k) - h_mid<2>(g, Eo, Ho, Po, i - 1, j, k)) * g.ax[0].inv_dd[i];
    } else {  // u = x, v = y: hv = Hy(p)-Hy(p-x), hu = Hx(p)-Hx(p-y)
        curl0 = (h_mid<1>(g, Eo, Ho, Po, i, j, k) - h_mid<1>(g, Eo, Ho, Po, i - 1, j, k)) * g.ax[0].inv_dd[i] -
                (h_mid<0>(g, Eo, Ho, Po, i, j, k) - h_mid<0>(g, Eo, Ho, Po, i, j - 1, k)) * g.ax[1].inv_dd[j];
    }
    // step n+1's E^{n+1} = ca E^n + cb curl^n (e_mid's update with the curl just formed: a
    // design edge is clear of the CPML, so no psi term enters), less its sources
    const float eo = Eo[fl];
    tape0[e] = make_float2(eo, curl0);
    const float2 cf = coef_at(g, fl);
    float e1 = cf.y != 0.f ? cf.x * eo + cf.y * curl0 : eo;
    for (uint32_t q = 0; q < f.n_src; q++)
        if (f.src[q].comp == c && f.src[q].id == id) e1 -= f.src[q].src * pl;
    const uint64_t sx = uint64_t(g.n1) * g.n2, sy = g.n2;
    float curl1;
    if (c == 0) {
        curl1 = (Hn[2 * nn + id] - Hn[2 * nn + id - sy]) * g.ax[1].inv_dd[j] - (Hn[nn + id] - Hn[nn + id - 1]) * g.ax[2].inv_dd[k];
    } else if (c == 1) {
        curl1 = (Hn[id] - Hn[id - 1]) * g.ax[2].inv_dd[k] - (Hn[2 * nn + id] - Hn[2 * nn + id - sx]) * g.ax[0].inv_dd[i];
    } else {
        curl1 = (Hn[nn + id] - Hn[nn + id - sx]) * g.ax[0].inv_dd[i] - (Hn[id] - Hn[id - sy]) * g.ax[1].inv_dd[j];
    }
    tape1[e] = make_float2(e1, curl1);
}

// ---- f64 forward (finite-difference checks) -------------------------------------------------

struct Geo64 {
    Geo g;                   // grid, axes (f32 tables, promoted), psi offsets
    const uint32_t *idx;     // one u32 set index per edge
    const double2 *sets;     // (ca, cb) in f64
    double k_mu;
};

template <int C>
__device__ __forceinline__ double h64(const Geo64 &G, const double *__restrict__ E, const double *__restrict__ H,
                                      double *__restrict__ psi, uint32_t i, uint32_t j, uint32_t k) {
    const Geo &g = G.g;
    constexpr int U = (C + 1) % 3, V = (C + 2) % 3;
    const uint64_t id = (uint64_t(i) * g.n1 + j) * g.n2 + k;
    const double hold = H[C * g.nn + id];
    const uint32_t pu = coord<U>(i, j, k), pv = coord<V>(i, j, k);
    if (pu + 1 >= dim<U>(g) || pv + 1 >= dim<V>(g)) return hold;
    const double ev = E[V * g.nn + id + stride<U>(g)] - E[V * g.nn + id];
    const double eu = E[U * g.nn + id + stride<V>(g)] - E[U * g.nn + id];
    double t1 = ev * double(g.ax[U].inv_d[pu]), t2 = eu * double(g.ax[V].inv_d[pv]);
    if (g.pml) {
        const int su = g.ax[U].slot_h[pu], sv = g.ax[V].slot_h[pv];
        if (su >= 0) {
            double &r = psi[psi_at<U>(g, 1, C, i, j, k, uint32_t(su))];
            r = double(g.ax[U].bh[pu]) * r + double(g.ax[U].ch[pu]) * ev * double(g.ax[U].raw_d[pu]);
            t1 += r;
        }
        if (sv >= 0) {
            double &r = psi[psi_at<V>(g, 1, C, i, j, k, uint32_t(sv))];
            r = double(g.ax[V].bh[pv]) * r + double(g.ax[V].ch[pv]) * eu * double(g.ax[V].raw_d[pv]);
            t2 += r;
        }
    }
    return hold - G.k_mu * (t1 - t2);
}

__global__ __launch_bounds__(256) void k_h64(Geo64 G, const double *__restrict__ E, double *__restrict__ H,
                                             double *__restrict__ psi) {
    const Geo &g = G.g;
    const uint32_t k = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, i = blockIdx.z;
    if (k >= g.n2 || j >= g.n1) return;
    const uint64_t id = (uint64_t(i) * g.n1 + j) * g.n2 + k;
    const double hx = h64<0>(G, E, H, psi, i, j, k), hy = h64<1>(G, E, H, psi, i, j, k), hz = h64<2>(G, E, H, psi, i, j, k);
    H[id] = hx, H[g.nn + id] = hy, H[2 * g.nn + id] = hz;
}

template <int C>
__device__ __forceinline__ double e64(const Geo64 &G, double eold, double hv, double hu, uint32_t i, uint32_t j,
                                      uint32_t k, uint64_t id, double *__restrict__ psi) {
    const Geo &g = G.g;
    constexpr int U = (C + 1) % 3, V = (C + 2) % 3;
    const double2 cf = G.sets[G.idx[C * g.nn + id]];
    if (cf.y == 0.0) return eold;
    const uint32_t pu = coord<U>(i, j, k), pv = coord<V>(i, j, k);
    double t1 = hv * double(g.ax[U].inv_dd[pu]), t2 = hu * double(g.ax[V].inv_dd[pv]);
    if (g.pml) {
        const int su = g.ax[U].slot_e[pu], sv = g.ax[V].slot_e[pv];
        if (su >= 0) {
            double &r = psi[psi_at<U>(g, 0, C, i, j, k, uint32_t(su))];
            r = double(g.ax[U].be[pu]) * r + double(g.ax[U].ce[pu]) * hv * double(g.ax[U].raw_dd[pu]);
            t1 += r;
        }
        if (sv >= 0) {
            double &r = psi[psi_at<V>(g, 0, C, i, j, k, uint32_t(sv))];
            r = double(g.ax[V].be[pv]) * r + double(g.ax[V].ce[pv]) * hu * double(g.ax[V].raw_dd[pv]);
            t2 += r;
        }
    }
    return cf.x * eold + cf.y * (t1 - t2);
}

__global__ __launch_bounds__(256) void k_e64(Geo64 G, double *__restrict__ E, const double *__restrict__ H,
                                             double *__restrict__ psi) {
    const Geo &g = G.g;
    const uint32_t k = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, i = blockIdx.z;
    if (k >= g.n2 || j >= g.n1) return;
    const uint64_t nn = g.nn, id = (uint64_t(i) * g.n1 + j) * g.n2 + k;
    const uint64_t im = i ? id - stride<0>(g) : id, jm = j ? id - stride<1>(g) : id, km = k ? id - 1 : id;
    const double hx = H[id], hy = H[nn + id], hz = H[2 * nn + id];
    const double ex = e64<0>(G, E[id], hz - H[2 * nn + jm], hy - H[nn + km], i, j, k, id, psi);
    const double ey = e64<1>(G, E[nn + id], hx - H[km], hz - H[2 * nn + im], i, j, k, id, psi);
    const double ez = e64<2>(G, E[2 * nn + id], hy - H[nn + im], hx - H[jm], i, j, k, id, psi);
    E[id] = ex, E[nn + id] = ey, E[2 * nn + id] = ez;
}

// sources of the driven port (cpu.rs:291-297) with the pulse (sim.rs:646-657) in f64, then the
// inductors (cpu.rs:298-303; one thread, in order: they may share an edge with a source), then
// the observations of this step
__global__ void k_fix64(uint64_t nn, const fg_source *__restrict__ src, uint32_t n_src, double *__restrict__ E,
                        uint64_t n, double dt, double t0, double tau, double w0, const uint64_t *__restrict__ obs,
                        uint32_t n_obs, double *__restrict__ series, const fg_inductor *__restrict__ ind,
                        uint32_t n_ind, double *__restrict__ cur) {
    const uint32_t q = threadIdx.x;
    const double t = (double(n) + 0.5) * dt, x = (t - t0) / tau;
    const double pl = exp(-x * x) * cos(w0 * (t - t0));
    for (uint32_t s = q; s < n_src; s += blockDim.x) E[src[s].comp * nn + src[s].id] -= double(src[s].src) * pl;
    __syncthreads();
    if (q == 0)
        for (uint32_t r = 0; r < n_ind; r++) {
            const uint64_t f = ind[r].comp * nn + ind[r].id;
            const double v = E[f] - double(ind[r].k_e) * cur[r];
            E[f] = v;
            cur[r] += double(ind[r].k_i) * v;
        }
    __syncthreads();
    for (uint32_t o = q; o < n_obs; o += blockDim.x) series[n * n_obs + o] = E[obs[o]];
}

// The f64 forward's thin wires (cpu.rs:270-273): H[h] += sum_t c_t E[e_t] after the H update,
// E being the step's input; each H node has one entry
__global__ __launch_bounds__(256) void k_wires64(const fg_wire_h *__restrict__ w, uint64_t n, const double *__restrict__ E,
                                                 double *__restrict__ H) {
    const uint64_t q = uint64_t(blockIdx.x) * 256 + threadIdx.x;
    if (q >= n) return;
    double s = 0.0;
    for (int t = 0; t < 4; t++) s += double(w[q].c[t]) * E[w[q].e[t]];
    H[w[q].h] += s;
}

// The f64 forward's Debye ADE (cpu.rs:239-260): E_old saved before the E update, then after it
// next = E - cb sum_k (1+ka_k)/2 J_k and J_k <- ka_k J_k + dd kb_k (next - E_old), with the
// problem's pole table and the edge's cb in f64
__global__ __launch_bounds__(256) void k_debye64(Geo64 G, FixArgs f, double *__restrict__ E, double *__restrict__ J,
                                                 double *__restrict__ old, int save) {
    const uint64_t d = uint64_t(blockIdx.x) * 256 + threadIdx.x;
    if (d >= f.n_debye) return;
    const fg_debye_edge e = f.debye[d];
    if (save) {
        old[d] = E[e.f];
        return;
    }
    const float *tab = f.dtab + size_t(e.set) * f.np * 2;
    const double cb = G.sets[G.idx[e.f]].y;
    double s0 = 0.0;
    for (uint32_t k = 0; k < f.np; k++) s0 += 0.5 * (1.0 + double(tab[2 * k])) * J[k * f.n_debye + d];
    const double next = E[e.f] - cb * s0;
    E[e.f] = next;
    const double de = next - old[d];
    for (uint32_t k = 0; k < f.np; k++)
        J[k * f.n_debye + d] = double(tab[2 * k]) * J[k * f.n_debye + d] + double(e.dd) * double(tab[2 * k + 1]) * de;
}

// ---- the sparse updates of the step, transposed -------------------------------------------
//
// The forward step ends (after its E update) with the Debye ADE and then the inductors
// (cpu.rs:239-260, 298-303); the adjoint step starts with their transposes, before the march:
//   inductors, last first: x = E - k_e I; E = x; I += k_i x  gives  lamE += k_i lamI; lamI -= k_e lamE;
//   Debye edge: next = E' - cb sum_k (1+ka_k)/2 J_k; J_k <- ka_k J_k + dd kb_k (next - E_old) gives
//     mu = lamE + sum_k dd kb_k lamJ_k (the adjoint of E'), lamJ_k <- ka_k lamJ_k - cb (1+ka_k)/2 mu,
//     and -sum_k dd kb_k lamJ_k (the adjoint of E_old = E^n), which k_adj_debye_post adds to the
//     march's output once the E update's transpose has turned mu into the rest of lamE^n.
// Inductor edges are lumped edges, which carry no Debye medium, so the two sets are disjoint.
__global__ __launch_bounds__(256) void k_adj_media_pre(Geo g, FixArgs f, float *__restrict__ lamE,
                                                       float *__restrict__ lamJ, float *__restrict__ lamI,
                                                       float *__restrict__ corr) {
    const uint64_t d = uint64_t(blockIdx.x) * 256 + threadIdx.x;
    if (d == 0)
        for (uint32_t q = f.n_ind; q-- > 0;) {
            const uint64_t fl = f.ind[q].comp * g.nn + f.ind[q].id;
            lamE[fl] += f.ind[q].k_i * lamI[q];
            lamI[q] -= f.ind[q].k_e * lamE[fl];
        }
    if (d >= f.n_debye) return;
    const fg_debye_edge e = f.debye[d];
    const float *tab = f.dtab + size_t(e.set) * f.np * 2;
    const float cb = coef_at(g, e.f).y;
    float back = 0.f;
    for (uint32_t k = 0; k < f.np; k++) back += e.dd * tab[2 * k + 1] * lamJ[k * f.n_debye + d];
    const float mu = lamE[e.f] + back;
    corr[d] = back;
    for (uint32_t k = 0; k < f.np; k++)
        lamJ[k * f.n_debye + d] = tab[2 * k] * lamJ[k * f.n_debye + d] - cb * (0.5f * (1.0f + tab[2 * k])) * mu;
    lamE[e.f] = mu;
}

__global__ __launch_bounds__(256) void k_adj_debye_post(FixArgs f, float *__restrict__ lamE, const float *__restrict__ corr) {
    const uint64_t d = uint64_t(blockIdx.x) * 256 + threadIdx.x;
    if (d < f.n_debye) lamE[f.debye[d].f] -= corr[d];
}

// The thin wires' transpose (cpu.rs:270-273): H'' = H' + W E^n gives lamE^n[e_t] += c_t lamH[h],
// lamH being the adjoint of the corrected H (the march's new lamH). One thread, since the
// terms of several H nodes may name one E edge.
__global__ void k_adj_wires_post(const fg_wire_h *__restrict__ w, uint64_t n, float *__restrict__ lamE,
                                 const float *__restrict__ lamH) {
    if (threadIdx.x || blockIdx.x) return;
    for (uint64_t q = 0; q < n; q++) {
        const float l = lamH[w[q].h];
        for (int t = 0; t < 4; t++) lamE[w[q].e[t]] += w[q].c[t] * l;
    }
}

}  // namespace fg

// ---- host ------------------------------------------------------------------------------------

struct fg_adjoint {
    fg_device *dev = nullptr;
    fg_sim_desc desc{};
    std::vector<float> eps;  // owned copy of the cell eps_r the desc points at
    std::vector<fg_point_probe> obs;
    fg_design design{};
    uint32_t driven = 0;
    fg_pulse pulse{};
    uint64_t box_cells = 0, steps = 0;
    uint32_t every = 100;
    fg_sim *sim = nullptr;
    fg_problem problem{};
    std::unique_ptr<fg::Session> s;
    std::vector<fg::DesignEdge> edges;
    bool built = false, have_forward = false;
    fg::DevBuf<float> snaps, lamE, lamA, lamH, lamH2, lamP[2];
    fg::DevBuf<uint64_t> d_edges, d_obs;
    fg::DevBuf<int32_t> dmap;
    fg::DevBuf<float2> tape[2];  // design-edge tapes of two segments (recompute k-1 while the adjoint reads k)
    bool whole = false;          // fg_design.whole_tape: the forward tapes every step, no recompute
    fg::DevBuf<float2> tape_all;  // with `whole`: the tape of every step, step n at n * design edges
    fg::DevBuf<double> gca, gcb, dseries;
    uint64_t n_snaps = 0, snap_len = 0;
    double ms[4] = {0, 0, 0, 0};
    hipStream_t astream = nullptr;  // the adjoint sweeps' stream (the session stream recomputes)
    // the x/y-slab class of a two-step adjoint pass runs on xy_stream, forked from and joined
    // back to the adjoint stream, next to the other classes (each is about one round of blocks)
    hipStream_t xy_stream = nullptr;
    hipEvent_t xy_fork = nullptr, xy_join = nullptr;
    // block lists of the two-step adjoint pass (k_adj_march2), per CPML class, for segments of
    // adj_seg planes
    fg::DevBuf<fg::MarchBlock> ablk2[3];
    uint32_t adj_seg = 0;
    // the step's sparse state (Debye pole states, inductor currents): its length in the
    // snapshots, and its adjoint (lamJ, lamI) with the Debye edges' E_old terms (corr)
    uint64_t dlen = 0, ilen = 0;
    fg::DevBuf<float> lamJ, lamI, corr;

    ~fg_adjoint() {
        if (astream) (void)hipStreamSynchronize(astream), (void)hipStreamDestroy(astream);
        if (xy_stream) (void)hipStreamSynchronize(xy_stream), (void)hipStreamDestroy(xy_stream);
        for (hipEvent_t e : {xy_fork, xy_join})
            if (e) (void)hipEventDestroy(e);
        s.reset();
        if (sim) fg_sim_free(sim);
    }
};

namespace fg {

static uint64_t box_index(const fg_design &d, uint64_t i, uint64_t j, uint64_t k) {
    return ((i - d.lo[0]) * (d.hi[1] - d.lo[1]) + (j - d.lo[1])) * (d.hi[2] - d.lo[2]) + (k - d.lo[2]);
}

// (Re)build the problem from the desc with the current eps copy, and its fused session.
static void rebuild(fg_adjoint &A) {
    A.s.reset();
    if (A.sim) fg_sim_free(A.sim);
    A.sim = nullptr;
    if (fg_sim_new(&A.desc, &A.sim) != FG_OK) throw HipError{g_error, FG_INVALID};
    uint64_t counts[4];
    fg_sim_counts(A.sim, counts);
    // resistive sheets weigh their faces by the H^2 the run has accumulated (cpu.rs:169-193), so a
    // step with sheets is not linear in the fields and has no transpose to run backward
    if (counts[2])
        throw HipError{"the adjoint covers curl updates, CPML, resistive ports, Debye media, inductors and thin "
                       "wires; resistive sheets make the step nonlinear in the fields",
                       FG_UNSUPPORTED};
    if (fg_sim_pack(A.sim, 0, &A.problem) != FG_OK) throw HipError{g_error, FG_INVALID};
    fg_extras x{};
    x.max_steps = A.steps;
    x.decay_db = INFINITY;
    x.engine = FG_ENGINE_FUSED;
    x.plane_k = -1;
    x.n_probes = uint32_t(A.obs.size());
    x.probes = A.obs.data();
    A.s.reset(session_new(A.dev, &A.problem, A.driven, &A.pulse, &x));
    A.dlen = A.s->fix.n_debye ? uint64_t(A.s->fix.np) * A.s->fix.n_debye : 0;
    A.ilen = A.s->fix.n_ind;
    if (A.edges.empty()) {
        const std::string err = design_map(*A.sim, A.desc, A.design.lo, A.design.hi, A.edges);
        if (!err.empty()) throw HipError{err, FG_INVALID};
    }
    A.built = true;
}

static void set_box(fg_adjoint &A, const float *eps_box) {
    const uint64_t n1 = A.desc.len[1] - 1, n2 = A.desc.len[2] - 1;
    const fg_design &d = A.design;
    bool changed = false;
    for (uint64_t i = d.lo[0]; i < d.hi[0]; i++)
        for (uint64_t j = d.lo[1]; j < d.hi[1]; j++)
            for (uint64_t k = d.lo[2]; k < d.hi[2]; k++) {
                float &v = A.eps[(i * n1 + j) * n2 + k];
                const float w = eps_box[box_index(d, i, j, k)];
                changed |= v != w;
                v = w;
            }
    if (changed) A.built = false;
}

static Geo geo(const fg_adjoint &A) { return A.s->geo; }

// The design edges' flat indices, the observations' flat indices and the map from (component,
// box node) to design edge, on the device (once).
static void design_tables(fg_adjoint &A) {
    if (A.d_edges.p) return;
    Session &s = *A.s;
    hipStream_t st = s.stream;
    const uint64_t nn = s.nn, ne = A.edges.size(), n_obs = A.obs.size();
    std::vector<uint64_t> f(ne), o(n_obs);
    for (uint64_t e = 0; e < ne; e++) f[e] = A.edges[e].f;
    for (uint64_t q = 0; q < n_obs; q++)
        o[q] = A.obs[q].field * nn + (uint64_t(A.obs[q].i) * s.n[1] + A.obs[q].j) * s.n[2] + A.obs[q].k;
    std::vector<uint64_t> fe(f);
    if (fe.empty()) fe.push_back(0);  // a box with no design edge: the tables exist, ne stays 0
    A.d_edges.upload(fe.data(), fe.size(), st);
    A.d_obs.upload(o.data(), o.size(), st);
    // design edge per (component, node of the box's nodes lo..hi)
    const fg_design &d = A.design;
    uint64_t nb = 1;
    for (int a = 0; a < 3; a++) nb *= d.hi[a] - d.lo[a] + 1;
    std::vector<int32_t> map(3 * nb, -1);
    for (uint64_t e = 0; e < ne; e++) {
        const uint64_t c = f[e] / nn, id = f[e] % nn;
        const uint64_t i = id / (uint64_t(s.n[1]) * s.n[2]), j = (id / s.n[2]) % s.n[1], k = id % s.n[2];
        const uint64_t b = ((i - d.lo[0]) * (d.hi[1] - d.lo[1] + 1) + (j - d.lo[1])) * (d.hi[2] - d.lo[2] + 1) + (k - d.lo[2]);
        map[c * nb + b] = int32_t(e);
    }
    A.dmap.upload(map.data(), map.size(), st);
    FG_CHECK(hipStreamSynchronize(st));  // the host vectors go out of scope
}

// The forward steps from the state in copy 0 at step `start` to step `end`, two per launch
// where the session allows it, taping (E^n, curl^n) on the design edges of each step into
// tape + (n - start) * design edges: k_adj_tape2 for a two-step launch, k_adj_tape for one.
// tune (the whole-tape forward): the segment tuner times its first taped passes.
static void steps_with_tape(fg_adjoint &A, uint64_t start, uint64_t end, float2 *tape, bool tune = false) {
    Session &s = *A.s;
    hipStream_t st = s.stream;
    const uint64_t ne = A.edges.size();
    const uint32_t eb = uint32_t((ne + 255) / 256);
    const Geo g = A.s->geo;
    s.cur = 0;
    s.steps = start;
    set_step_counter(s);
    uint64_t n = start;
    auto advance = [&](uint64_t upto) {
        for (; n < upto;) {
            const int prev = s.cur;
            if (s.fuse2 && n + 2 <= end) {
                launch_pair(s, uint32_t(n - start));
                if (ne)
                    hipLaunchKernelGGL(k_adj_tape2, dim3(eb), dim3(256), 0, st, g, s.fix, A.d_edges.p, uint32_t(ne),
                                       s.E[prev].p, s.H[prev].p, s.psi[prev].p, s.H[s.cur].p, uint32_t(n - start),
                                       tape + (n - start) * ne, tape + (n + 1 - start) * ne);
                n += 2;
            } else {
                launch_step(s, uint32_t(n - start), n);
                if (ne)
                    hipLaunchKernelGGL(k_adj_tape, dim3(eb), dim3(256), 0, st, g, A.d_edges.p, uint32_t(ne), s.E[prev].p,
                                       s.H[s.cur].p, tape + (n - start) * ne);
                n += 1;
            }
        }
    };
    // the whole-tape forward (one call over every step) times the two-step march's segment
    // lengths on its first taped passes, as enqueue_steps does for a plain run
    if (tune && s.fuse2 && !s.seg2_try.empty() && end - start >= seg2_tune_steps(s))
        tune_seg2(s, [&](uint64_t k) { advance(n + k); });
    advance(end);
    s.steps = end;
}

static void forward(fg_adjoint &A, float *series) {
    if (!A.built) rebuild(A);
    Session &s = *A.s;
    hipStream_t st = s.stream;
    reset(s);
    s.observe = true;
    if (A.whole) {
        // the tape of every step, no snapshots: the backward recomputes nothing
        design_tables(A);
        const uint64_t ne = A.edges.size(), len = std::max<uint64_t>(A.steps * ne, 1);
        if (A.tape_all.cap < len) {
            size_t free_b = 0, total_b = 0;
            FG_CHECK(hipMemGetInfo(&free_b, &total_b));
            if (len * 8 > free_b)
                throw HipError{"the whole tape (" + std::to_string((len * 8) >> 20) + " MiB) does not fit in this GPU",
                               FG_TOO_LARGE};
            A.tape_all.alloc(len);
        }
        A.n_snaps = 0;
        const auto t0 = std::chrono::steady_clock::now();
        steps_with_tape(A, 0, A.steps, A.tape_all.p, true);
        FG_CHECK(hipGetLastError());
        FG_CHECK(hipStreamSynchronize(st));
        A.ms[0] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        if (series && A.steps && !A.obs.empty())
            FG_CHECK(hipMemcpy(series, s.probes.p, A.steps * A.obs.size() * 4, hipMemcpyDeviceToHost));
        A.have_forward = true;
        return;
    }
    const uint64_t psi_n = s.psi[0].n, nn = s.nn;
    A.snap_len = 6 * nn + psi_n + A.dlen + A.ilen;
    A.n_snaps = (A.steps + A.every - 1) / A.every;  // segment k starts at k*every; k = 0 is all zeros
    if (A.snaps.n < (A.n_snaps ? A.n_snaps - 1 : 0) * A.snap_len) {
        size_t free_b = 0, total_b = 0;
        FG_CHECK(hipMemGetInfo(&free_b, &total_b));
        const uint64_t need = (A.n_snaps - 1) * A.snap_len * 4;
        if (need > free_b)
            throw HipError{"the field snapshots (" + std::to_string(need >> 20) + " MiB) do not fit in this GPU", FG_TOO_LARGE};
        A.snaps.alloc(std::max<uint64_t>((A.n_snaps - 1) * A.snap_len, 1));
    }
    const auto t0 = std::chrono::steady_clock::now();
    for (uint64_t k = 0; k < A.n_snaps; k++) {
        if (k) {  // snapshot at the start of segment k
            float *d = A.snaps.p + (k - 1) * A.snap_len;
            FG_CHECK(hipMemcpyAsync(d, s.E[s.cur].p, 3 * nn * 4, hipMemcpyDeviceToDevice, st));
            FG_CHECK(hipMemcpyAsync(d + 3 * nn, s.H[s.cur].p, 3 * nn * 4, hipMemcpyDeviceToDevice, st));
            if (psi_n) FG_CHECK(hipMemcpyAsync(d + 6 * nn, s.psi[s.cur].p, psi_n * 4, hipMemcpyDeviceToDevice, st));
            if (A.dlen) FG_CHECK(hipMemcpyAsync(d + 6 * nn + psi_n, s.dstate.p, A.dlen * 4, hipMemcpyDeviceToDevice, st));
            if (A.ilen)
                FG_CHECK(hipMemcpyAsync(d + 6 * nn + psi_n + A.dlen, s.ind_i.p, A.ilen * 4, hipMemcpyDeviceToDevice, st));
        }
        enqueue_steps(s, std::min<uint64_t>(A.every, A.steps - k * A.every));
    }
    FG_CHECK(hipStreamSynchronize(st));
    A.ms[0] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (series && A.steps && !A.obs.empty())
        FG_CHECK(hipMemcpy(series, s.probes.p, A.steps * A.obs.size() * 4, hipMemcpyDeviceToHost));
    A.have_forward = true;
}

// One adjoint step through k_adj_march over the session's march block lists (the same reach as
// the forward march: one halo row/lane each side and one warm-up plane).
static void adj_march(Session &s, hipStream_t st, const float *Lo, float *Ln, const float *Mo, float *Mn,
                      const float *po, float *pn, const GradArgs &ga, const InjectArgs &in) {
    const Geo &g = s.geo;
    const dim3 block(64, s.tj + 2);
    const uint32_t ni = uint32_t(s.blk_int.n), nz = uint32_t(s.blk_z.n), np_ = uint32_t(s.blk_pml.n);
    const MarchBlock *bi = s.blk_int.p, *bz = s.blk_z.p, *bp = s.blk_pml.p;
    const hipStream_t cs[3] = {st, st, st};
#define FG_AM2(TJ, W, L)                                                                                                     \
    do {                                                                                                                     \
        if (ni) hipLaunchKernelGGL((k_adj_march<TJ, 0, W, L>), dim3(ni), block, 0, cs[0], g, bi, Lo, Ln, Mo, Mn, po, pn, s.seg, ga, in); \
        if (nz) hipLaunchKernelGGL((k_adj_march<TJ, 1, W, L>), dim3(nz), block, 0, cs[1], g, bz, Lo, Ln, Mo, Mn, po, pn, s.seg, ga, in); \
        if (np_) hipLaunchKernelGGL((k_adj_march<TJ, 2, W, L>), dim3(np_), block, 0, cs[2], g, bp, Lo, Ln, Mo, Mn, po, pn, s.seg, ga, in); \
    } while (0)
#define FG_AM(TJ)                                                                     \
    do {                                                                              \
        const bool lds = g.n_sets <= uint32_t(SETS_LDS);                              \
        if (g.wide) {                                                                 \
            if (lds) FG_AM2(TJ, true, true); else FG_AM2(TJ, true, false);             \
        } else {                                                                      \
            if (lds) FG_AM2(TJ, false, true); else FG_AM2(TJ, false, false);           \
        }                                                                             \
    } while (0)
    if (s.tj == 8) FG_AM(8);
    else FG_AM(14);
#undef FG_AM
#undef FG_AM2
}

// Rows per block of the two-step adjoint's x/y-slab class (the other classes take 12).
constexpr uint32_t ADJ_TJ_XY = 6;

// Two adjoint steps through k_adj_march2 over the two-step adjoint block lists (march2's tile,
// reading one plane further along i), one launch per CPML class: the x/y-slab class on
// xy_stream next to the others (FIELDGRID_ADJ_STREAMS=0: all on the stream, one after another).
static void adj_march2(fg_adjoint &A, hipStream_t st, const float *Lo, float *Ln, const float *Mo, float *Mn,
                       const float *po, float *pn, const GradArgs &ga, const float2 *tapeB, const InjectArgs &in,
                       const double *dsB) {
    Session &s = *A.s;
    const Geo &g = s.geo;
    if (!A.adj_seg) {
        const char *e = getenv("FIELDGRID_ADJ_SEG2");
        const uint32_t sg = e ? uint32_t(std::max(1, std::min(atoi(e), SEG_MAX))) : s.seg2_base;
        std::vector<MarchBlock> cls[3];
        march_classes2(s, 12, sg, cls, 1);
        // The x/y-slab class runs 6-row tiles (each 12-row tile as two): at 12 rows its 1024
        // threads leave 128 registers a lane and it spilled 45; 640 threads leave 170 and it
        // holds 160. Every node's terms are the same whatever the tile, so the results are too.
        {
            std::vector<MarchBlock> half;
            for (const MarchBlock &m : cls[2])
                for (uint32_t h = 0; h < 2; h++)
                    if ((2 * m.by + h) * ADJ_TJ_XY < s.n[1]) half.push_back(MarchBlock{m.bx, 2 * m.by + h, m.bz});
            cls[2].swap(half);
        }
        for (int c = 0; c < 3; c++) {
            xcd_order(cls[c], s.xcds, MarchBlock{~0u, ~0u, ~0u});
            if (cls[c].empty()) A.ablk2[c].alloc(1);
            else A.ablk2[c].upload(cls[c].data(), cls[c].size(), s.stream);
            A.ablk2[c].n = cls[c].size();
        }
        FG_CHECK(hipStreamSynchronize(s.stream));
        A.adj_seg = sg;
        const char *cs = getenv("FIELDGRID_ADJ_STREAMS");
        if (!A.xy_stream && !(cs && cs[0] == '0') && !cls[2].empty() && !(cls[0].empty() && cls[1].empty())) {
            FG_CHECK(hipStreamCreateWithFlags(&A.xy_stream, hipStreamNonBlocking));
            FG_CHECK(hipEventCreateWithFlags(&A.xy_fork, hipEventDisableTiming));
            FG_CHECK(hipEventCreateWithFlags(&A.xy_join, hipEventDisableTiming));
        }
    }
    const dim3 block(64, 12 + 4), block_xy(64, ADJ_TJ_XY + 4);
    const uint32_t n0 = uint32_t(A.ablk2[0].n), n1 = uint32_t(A.ablk2[1].n), n2 = uint32_t(A.ablk2[2].n);
    const MarchBlock *b0 = A.ablk2[0].p, *b1 = A.ablk2[1].p, *b2 = A.ablk2[2].p;
    const uint32_t sg = A.adj_seg;
    // every node is owned by one block of one class, so the classes may run in any order
    hipStream_t xs = st;
    if (A.xy_stream && n2) {
        FG_CHECK(hipEventRecord(A.xy_fork, st));
        FG_CHECK(hipStreamWaitEvent(A.xy_stream, A.xy_fork, 0));
        xs = A.xy_stream;
    }
#define FG_AM2(W, L)                                                                                             \
    do {                                                                                                         \
        if (n2) hipLaunchKernelGGL((k_adj_march2<ADJ_TJ_XY, 2, W, L>), dim3(n2), block_xy, 0, xs, g, b2, Lo, Ln, Mo, \
                                   Mn, po, pn, sg, ga, tapeB, in, dsB);                                          \
        if (n1) hipLaunchKernelGGL((k_adj_march2<12, 1, W, L>), dim3(n1), block, 0, st, g, b1, Lo, Ln, Mo, Mn, po, pn, \
                                   sg, ga, tapeB, in, dsB);                                                      \
        if (n0) hipLaunchKernelGGL((k_adj_march2<12, 0, W, L>), dim3(n0), block, 0, st, g, b0, Lo, Ln, Mo, Mn, po, pn, \
                                   sg, ga, tapeB, in, dsB);                                                      \
    } while (0)
    const bool lds = g.n_sets <= uint32_t(SETS_LDS);
    if (g.wide) {
        if (lds) FG_AM2(true, true);
        else FG_AM2(true, false);
    } else {
        if (lds) FG_AM2(false, true);
        else FG_AM2(false, false);
    }
#undef FG_AM2
    if (xs != st) {
        FG_CHECK(hipEventRecord(A.xy_join, xs));
        FG_CHECK(hipStreamWaitEvent(st, A.xy_join, 0));
    }
}

static void backward(fg_adjoint &A, const double *dseries, double *grad) {
    if (!A.have_forward) throw HipError{"fg_adjoint_backward needs a forward run first", FG_INVALID};
    Session &s = *A.s;
    hipStream_t st = s.stream;
    const uint64_t nn = s.nn, psi_n = s.psi[0].n, ne = A.edges.size(), n_obs = A.obs.size();
    const Geo g = geo(A);
    design_tables(A);
    if (A.lamE.n != 3 * nn) {
        A.lamE.alloc(3 * nn);
        A.lamA.alloc(3 * nn);
        A.lamH.alloc(3 * nn);
        A.lamH2.alloc(3 * nn);
        A.lamP[0].alloc(std::max<uint64_t>(psi_n, 1));
        A.lamP[1].alloc(std::max<uint64_t>(psi_n, 1));
        A.gca.alloc(std::max<uint64_t>(ne, 1));
        A.gcb.alloc(std::max<uint64_t>(ne, 1));
    }
    if (!A.whole && A.tape[0].cap < std::max<uint64_t>(A.every * ne, 1)) {
        A.tape[0].alloc(std::max<uint64_t>(A.every * ne, 1));
        A.tape[1].alloc(std::max<uint64_t>(A.every * ne, 1));
    }
    const bool media = A.dlen || A.ilen || s.wires.n;
    if (media) {
        if (A.lamJ.n != std::max<uint64_t>(A.dlen, 1)) A.lamJ.alloc(std::max<uint64_t>(A.dlen, 1));
        if (A.lamI.n != std::max<uint64_t>(A.ilen, 1)) A.lamI.alloc(std::max<uint64_t>(A.ilen, 1));
        if (A.corr.n != std::max<uint64_t>(s.fix.n_debye, 1)) A.corr.alloc(std::max<uint64_t>(s.fix.n_debye, 1));
        A.lamJ.zero(st), A.lamI.zero(st);
    }
    A.lamE.zero(st), A.lamA.zero(st), A.lamH.zero(st), A.lamH2.zero(st), A.lamP[0].zero(st), A.lamP[1].zero(st);
    // the fused adjoint march (one pass per adjoint step) unless FIELDGRID_ADJ_MARCH=0 asks for
    // the two gather kernels; the march ping-pongs lamE (lamE/lamA) and lamH (lamH/lamH2)
    const char *am = getenv("FIELDGRID_ADJ_MARCH");
    const bool march = s.fused && (!(am && am[0] == '0') || media);  // the gather kernels have no media
    // two adjoint steps per pass (k_adj_march2) unless FIELDGRID_ADJ_PAIRS=0 asks for one: on the
    // 192x192x112 design grid the backward takes 589 ms against 628 ms with one step per pass,
    // its recompute overlapped either way (the step is latency-bound, not bandwidth-bound: the
    // two-step tile computes more per owned node-step and each class is about one round of
    // blocks, so halving the passes gains little); runs with Debye media, inductors or thin
    // wires take one step per pass (their transposes sit between the steps)
    const char *ap = getenv("FIELDGRID_ADJ_PAIRS");
    const bool pairs = march && !media && !(ap && ap[0] == '0');
    // The recompute marches at the segment length set before any tuning: the forward's tuner
    // times the forward alone, and the recompute shares the GPU with the adjoint sweep (on the
    // 192x192x112 design grid the forward is fastest at shorter segments, the backward with a
    // recompute at 16 planes). The tuned length is restored at the end.
    const uint32_t tuned = s.seg2;
    if (!A.whole && s.fuse2 && tuned != s.seg2_base) {
        s.seg2 = s.seg2_base;
        plan_march2(s);
    }
    float *LE[2] = {A.lamE.p, A.lamA.p}, *LH[2] = {A.lamH.p, A.lamH2.p};
    int le = 0;
    A.gca.zero(st), A.gcb.zero(st);
    A.dseries.upload(dseries, A.steps * n_obs, st);
    // The session stream recomputes the segments (restore, forward with tape); the adjoint
    // stream sweeps them backward. Segment k-1 is recomputed while the adjoint sweeps segment k:
    // the tapes alternate between two buffers, and the recompute of segment k waits until the
    // adjoint of segment k+2 has read the buffer it overwrites. FIELDGRID_ADJ_OVERLAP=0 runs
    // both on the session stream.
    const char *ov = getenv("FIELDGRID_ADJ_OVERLAP");
    if (!A.astream) FG_CHECK(hipStreamCreateWithFlags(&A.astream, hipStreamNonBlocking));
    hipStream_t as = ov && ov[0] == '0' ? st : A.astream;
    // per segment: recompute start/end on st, adjoint start/end on as; plus the setup event.
    // With the whole tape there is one segment, all steps, and nothing to recompute.
    const uint64_t segs = A.whole ? 1 : A.n_snaps;
    std::vector<hipEvent_t> ev(4 * segs + 1);
    for (auto &e : ev) FG_CHECK(hipEventCreate(&e));
    hipEvent_t ready = ev[4 * segs];
    const dim3 block(64, 4), grid((g.n2 + 63) / 64, (g.n1 + 3) / 4, g.n0);
    GradArgs ga{};
    InjectArgs in{A.d_obs.p, nullptr, uint32_t(n_obs)};
    for (int a = 0; a < 3; a++) ga.lo[a] = A.design.lo[a], ga.nb[a] = A.design.hi[a] - A.design.lo[a] + 1;
    ga.map = A.dmap.p, ga.gca = A.gca.p, ga.gcb = A.gcb.p;
    int pin = 0;
    const auto t0 = std::chrono::steady_clock::now();
    FG_CHECK(hipEventRecord(ready, st));
    FG_CHECK(hipStreamWaitEvent(as, ready, 0));
    // dJ/ds of the last step; every earlier one enters at the end of the next step's transpose
    if (n_obs && A.steps)
        hipLaunchKernelGGL(k_adj_inject, dim3(1), dim3(64), 0, as, g, A.lamE.p, march ? nullptr : A.lamA.p, A.d_obs.p,
                           uint32_t(n_obs), A.dseries.p + (A.steps - 1) * n_obs);
    for (uint64_t k = segs; k-- > 0;) {
        const uint64_t start = A.whole ? 0 : k * A.every, end = A.whole ? A.steps : std::min(start + A.every, A.steps);
        float2 *tape = A.whole ? A.tape_all.p : A.tape[k & 1].p;
        if (k + 2 < segs) FG_CHECK(hipStreamWaitEvent(st, ev[4 * (k + 2) + 3], 0));
        FG_CHECK(hipEventRecord(ev[4 * k], st));
        if (!A.whole) {
            // restore the segment's starting fields into copy 0
            if (k == 0) {
                s.E[0].zero(st), s.H[0].zero(st), s.psi[0].zero(st);
                if (A.dlen) s.dstate.zero(st);
                if (A.ilen) FG_CHECK(hipMemcpyAsync(s.ind_i.p, s.ind_i0.data(), A.ilen * 4, hipMemcpyHostToDevice, st));
            } else {
                const float *src = A.snaps.p + (k - 1) * A.snap_len;
                FG_CHECK(hipMemcpyAsync(s.E[0].p, src, 3 * nn * 4, hipMemcpyDeviceToDevice, st));
                FG_CHECK(hipMemcpyAsync(s.H[0].p, src + 3 * nn, 3 * nn * 4, hipMemcpyDeviceToDevice, st));
                if (psi_n) FG_CHECK(hipMemcpyAsync(s.psi[0].p, src + 6 * nn, psi_n * 4, hipMemcpyDeviceToDevice, st));
                if (A.dlen)
                    FG_CHECK(hipMemcpyAsync(s.dstate.p, src + 6 * nn + psi_n, A.dlen * 4, hipMemcpyDeviceToDevice, st));
                if (A.ilen)
                    FG_CHECK(hipMemcpyAsync(s.ind_i.p, src + 6 * nn + psi_n + A.dlen, A.ilen * 4, hipMemcpyDeviceToDevice,
                                            st));
            }
            s.observe = false;  // the series of this segment is already known
            steps_with_tape(A, start, end, tape);
            s.observe = true;
        }
        FG_CHECK(hipEventRecord(ev[4 * k + 1], st));
        FG_CHECK(hipStreamWaitEvent(as, ev[4 * k + 1], 0));
        FG_CHECK(hipEventRecord(ev[4 * k + 2], as));
        for (uint64_t n = end; n-- > start;) {
            ga.tape = tape + (n - start) * ne;
            in.ds = n ? A.dseries.p + (n - 1) * n_obs : nullptr;
            if (pairs && n >= start + 1) {  // steps n and n-1 in one pass
                const uint64_t m = n - 1;
                adj_march2(A, as, LE[le], LE[le ^ 1], LH[le], LH[le ^ 1], A.lamP[pin].p, A.lamP[pin ^ 1].p, ga,
                           tape + (m - start) * ne, in, m ? A.dseries.p + (m - 1) * n_obs : nullptr);
                le ^= 1;
                pin ^= 1;
                n = m;  // the loop's decrement then moves past step n-1
                continue;
            }
            if (march) {
                const uint32_t mb = uint32_t((std::max<uint64_t>(s.fix.n_debye, 1) + 255) / 256);
                if (media)
                    hipLaunchKernelGGL(k_adj_media_pre, dim3(mb), dim3(256), 0, as, g, s.fix, LE[le], A.lamJ.p, A.lamI.p,
                                       A.corr.p);
                adj_march(s, as, LE[le], LE[le ^ 1], LH[le], LH[le ^ 1], A.lamP[pin].p, A.lamP[pin ^ 1].p, ga, in);
                if (s.fix.n_debye) hipLaunchKernelGGL(k_adj_debye_post, dim3(mb), dim3(256), 0, as, s.fix, LE[le ^ 1], A.corr.p);
                if (s.wires.n)
                    hipLaunchKernelGGL(k_adj_wires_post, dim3(1), dim3(64), 0, as, s.wires.p, uint64_t(s.wires.n), LE[le ^ 1],
                                       LH[le ^ 1]);
                le ^= 1;
            } else {
                hipLaunchKernelGGL(k_adj_e, grid, block, 0, as, g, A.lamA.p, A.lamE.p, A.lamH.p, A.lamP[pin].p, ga);
                hipLaunchKernelGGL(k_adj_h, grid, block, 0, as, g, A.lamE.p, A.lamA.p, A.lamH.p, A.lamP[pin].p,
                                   A.lamP[pin ^ 1].p, in);
            }
            pin ^= 1;
        }
        FG_CHECK(hipEventRecord(ev[4 * k + 3], as));
        FG_CHECK(hipGetLastError());
    }
    FG_CHECK(hipStreamSynchronize(as));
    FG_CHECK(hipStreamSynchronize(st));
    A.ms[1] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (s.seg2 != tuned) {
        s.seg2 = tuned;
        plan_march2(s);
        FG_CHECK(hipStreamSynchronize(st));
    }
    A.ms[2] = A.ms[3] = 0.0;  // busy time of each stream (they overlap)
    for (uint64_t k = 0; k < segs; k++) {
        float a = 0.f, b = 0.f;
        FG_CHECK(hipEventElapsedTime(&a, ev[4 * k], ev[4 * k + 1]));
        FG_CHECK(hipEventElapsedTime(&b, ev[4 * k + 2], ev[4 * k + 3]));
        A.ms[2] += a, A.ms[3] += b;
    }
    if (A.whole) A.ms[2] = 0.0;  // nothing recomputed (the events only bracket an empty interval)
    for (auto &e : ev) (void)hipEventDestroy(e);
    // chain rule per design edge: ca = (1-x)/(1+x), cb = dt/(eps (1+x)), x = sig dt / (2 eps)
    // (sim.rs:431-441), eps = EPS0 (fixed + sum w eps_r)
    std::vector<double> hca(ne), hcb(ne);
    if (ne) {
        FG_CHECK(hipMemcpy(hca.data(), A.gca.p, ne * 8, hipMemcpyDeviceToHost));
        FG_CHECK(hipMemcpy(hcb.data(), A.gcb.p, ne * 8, hipMemcpyDeviceToHost));
    }
    std::fill(grad, grad + A.box_cells, 0.0);
    const double dt = fg_sim_dt(A.sim), eps0 = 8.8541878128e-12;
    const uint64_t n1 = A.desc.len[1] - 1, n2 = A.desc.len[2] - 1;
    const fg_design &d = A.design;
    for (uint64_t e = 0; e < ne; e++) {
        const DesignEdge &de = A.edges[e];
        double er = de.fixed;
        for (int q = 0; q < 4; q++)
            if (de.cell[q] != UINT32_MAX) {
                const uint64_t bz = d.hi[2] - d.lo[2], by = d.hi[1] - d.lo[1];
                const uint64_t ci = de.cell[q] / (by * bz) + d.lo[0], cj = (de.cell[q] / bz) % by + d.lo[1],
                               ck = de.cell[q] % bz + d.lo[2];
                er += de.w[q] * double(A.eps[(ci * n1 + cj) * n2 + ck]);
            }
        const double eps = eps0 * er, x = de.sig * dt / (2.0 * eps);
        const double dca = 2.0 * x / (eps * (1.0 + x) * (1.0 + x)), dcb = -dt / (eps * eps * (1.0 + x) * (1.0 + x));
        const double gedge = hca[e] * dca + hcb[e] * dcb;
        for (int q = 0; q < 4; q++)
            if (de.cell[q] != UINT32_MAX) grad[de.cell[q]] += gedge * eps0 * de.w[q];
    }
}

static void forward64(fg_adjoint &A, const double *eps_box, double *series) {
    if (!A.built) rebuild(A);
    Session &s = *A.s;
    hipStream_t st = s.stream;
    const uint64_t nn = s.nn, psi_n = s.psi[0].n, n_obs = A.obs.size();
    const fg_problem &P = A.problem;
    // coefficient sets in f64: the problem's sets, then one pair per design edge from eps_box
    std::vector<double2> sets(P.n_sets);
    for (uint32_t q = 0; q < P.n_sets; q++) sets[q] = make_double2(P.coef_sets[2 * q], P.coef_sets[2 * q + 1]);
    std::vector<uint32_t> idx(3 * nn);
    const uint16_t *narrow = reinterpret_cast<const uint16_t *>(P.coef_index);
    for (uint64_t f = 0; f < 3 * nn; f++) idx[f] = P.coef_wide ? P.coef_index[f] : narrow[f];
    const double dt = fg_sim_dt(A.sim), eps0 = 8.8541878128e-12;
    const uint64_t n1 = A.desc.len[1] - 1, n2 = A.desc.len[2] - 1;
    const fg_design &d = A.design;
    const uint64_t bz = d.hi[2] - d.lo[2], by = d.hi[1] - d.lo[1];
    for (const DesignEdge &de : A.edges) {
        double er = de.fixed;
        for (int q = 0; q < 4; q++)
            if (de.cell[q] != UINT32_MAX) {
                const uint64_t ci = de.cell[q] / (by * bz) + d.lo[0], cj = (de.cell[q] / bz) % by + d.lo[1],
                               ck = de.cell[q] % bz + d.lo[2];
                er += de.w[q] * (eps_box ? eps_box[de.cell[q]] : double(A.eps[(ci * n1 + cj) * n2 + ck]));
            }
        const double eps = eps0 * er, x = de.sig * dt / (2.0 * eps);
        idx[de.f] = uint32_t(sets.size());
        sets.push_back(make_double2((1.0 - x) / (1.0 + x), dt / eps / (1.0 + x)));
    }
    DevBuf<double2> dsets;
    DevBuf<uint32_t> didx;
    DevBuf<double> E, H, psi, out;
    DevBuf<uint64_t> dobs;
    DevBuf<fg_source> src;
    dsets.upload(sets.data(), sets.size(), st);
    didx.upload(idx.data(), idx.size(), st);
    E.alloc(3 * nn), H.alloc(3 * nn), psi.alloc(std::max<uint64_t>(psi_n, 1)), out.alloc(std::max<uint64_t>(A.steps * n_obs, 1));
    E.zero(st), H.zero(st), psi.zero(st);
    std::vector<uint64_t> o(n_obs);
    for (uint64_t q = 0; q < n_obs; q++)
        o[q] = A.obs[q].field * nn + (uint64_t(A.obs[q].i) * s.n[1] + A.obs[q].j) * s.n[2] + A.obs[q].k;
    dobs.upload(o.data(), o.size(), st);
    uint32_t n_src = 0;
    if (P.n_ports) {
        const fg_port &dp = P.ports[A.driven];
        n_src = dp.n_sources;
        src.upload(P.sources + dp.first_source, n_src, st);
    }
    Geo64 G{s.geo, didx.p, dsets.p, dt / 1.25663706212e-6};
    const double tau = std::sqrt(std::log(10.0)) / (3.14159265358979323846 * A.pulse.fc), t0 = 4.0 * tau;
    const double w0 = 2.0 * 3.14159265358979323846 * A.pulse.f0;
    const dim3 block(64, 4), grid((s.n[2] + 63) / 64, (s.n[1] + 3) / 4, s.n[0]);
    // the step's sparse state in f64: Debye pole states and E_old, inductor currents
    const FixArgs &fx = s.fix;
    DevBuf<double> J, old, cur;
    J.alloc(std::max<uint64_t>(A.dlen, 1)), old.alloc(std::max<uint64_t>(fx.n_debye, 1));
    cur.alloc(std::max<uint64_t>(A.ilen, 1));
    J.zero(st), cur.zero(st);
    const uint32_t db = uint32_t((std::max<uint64_t>(fx.n_debye, 1) + 255) / 256);
    const uint32_t wb = uint32_t((std::max<uint64_t>(s.wires.n, 1) + 255) / 256);
    for (uint64_t n = 0; n < A.steps; n++) {
        hipLaunchKernelGGL(k_h64, grid, block, 0, st, G, E.p, H.p, psi.p);
        if (s.wires.n) hipLaunchKernelGGL(k_wires64, dim3(wb), dim3(256), 0, st, s.wires.p, uint64_t(s.wires.n), E.p, H.p);
        if (fx.n_debye) hipLaunchKernelGGL(k_debye64, dim3(db), dim3(256), 0, st, G, fx, E.p, J.p, old.p, 1);
        hipLaunchKernelGGL(k_e64, grid, block, 0, st, G, E.p, H.p, psi.p);
        if (fx.n_debye) hipLaunchKernelGGL(k_debye64, dim3(db), dim3(256), 0, st, G, fx, E.p, J.p, old.p, 0);
        hipLaunchKernelGGL(k_fix64, dim3(1), dim3(64), 0, st, nn, src.p, n_src, E.p, n, dt, t0, tau, w0, dobs.p,
                           uint32_t(n_obs), out.p, fx.ind, fx.n_ind, cur.p);
    }
    FG_CHECK(hipGetLastError());
    if (series && A.steps * n_obs)
        FG_CHECK(hipMemcpyAsync(series, out.p, A.steps * n_obs * 8, hipMemcpyDeviceToHost, st));
    FG_CHECK(hipStreamSynchronize(st));
}

}  // namespace fg

extern "C" {

int fg_adjoint_new(fg_device *dev, const fg_sim_desc *desc, const fg_design *design, uint32_t driven,
                   const fg_pulse *pulse, fg_adjoint **out) {
    using namespace fg;
    return guarded([&] {
        if (!dev || !desc || !design || !pulse || !out) throw HipError{"fg_adjoint_new: NULL argument", FG_INVALID};
        auto A = std::make_unique<fg_adjoint>();
        A->dev = dev;
        A->desc = *desc;
        for (int a = 0; a < 3; a++)
            if (!desc->line[a] || desc->len[a] < 2) throw HipError{"a grid line needs at least two nodes", FG_INVALID};
        const uint64_t cells = uint64_t(desc->len[0] - 1) * (desc->len[1] - 1) * (desc->len[2] - 1);
        A->eps.assign(cells, 1.0f);
        if (desc->eps) std::memcpy(A->eps.data(), desc->eps, cells * 4);
        A->desc.eps = A->eps.data();
        A->design = *design;
        A->obs.assign(design->obs, design->obs + design->n_obs);
        for (const fg_point_probe &o : A->obs)
            if (o.field > 2 || o.i >= desc->len[0] || o.j >= desc->len[1] || o.k >= desc->len[2])
                throw HipError{"an observation is not an E component inside the grid", FG_INVALID};
        for (int a = 0; a < 3; a++)
            if (design->lo[a] >= design->hi[a] || design->hi[a] > desc->len[a] - 1)
                throw HipError{"the design box lies outside the grid", FG_INVALID};
        A->design.obs = A->obs.data();
        A->whole = design->whole_tape != 0;
        A->box_cells = uint64_t(design->hi[0] - design->lo[0]) * (design->hi[1] - design->lo[1]) * (design->hi[2] - design->lo[2]);
        A->steps = design->steps;
        A->every = design->every ? design->every : 100;
        A->driven = driven;
        A->pulse = *pulse;
        FG_CHECK(hipSetDevice(dev->ordinal));
        rebuild(*A);
        *out = A.release();
        return int(FG_OK);
    });
}

int fg_adjoint_forward(fg_adjoint *adj, const float *eps_box, float *series) {
    using namespace fg;
    return guarded([&] {
        if (!adj) throw HipError{"fg_adjoint_forward: NULL handle", FG_INVALID};
        FG_CHECK(hipSetDevice(adj->dev->ordinal));
        if (eps_box) set_box(*adj, eps_box);
        forward(*adj, series);
        return int(FG_OK);
    });
}

int fg_adjoint_backward(fg_adjoint *adj, const double *dseries, double *grad) {
    using namespace fg;
    return guarded([&] {
        if (!adj || !dseries || !grad) throw HipError{"fg_adjoint_backward: NULL argument", FG_INVALID};
        FG_CHECK(hipSetDevice(adj->dev->ordinal));
        backward(*adj, dseries, grad);
        return int(FG_OK);
    });
}

int fg_adjoint_forward64(fg_adjoint *adj, const double *eps_box, double *series) {
    using namespace fg;
    return guarded([&] {
        if (!adj) throw HipError{"fg_adjoint_forward64: NULL handle", FG_INVALID};
        FG_CHECK(hipSetDevice(adj->dev->ordinal));
        forward64(*adj, eps_box, series);
        return int(FG_OK);
    });
}

int fg_adjoint_timing(const fg_adjoint *adj, double ms[4], uint64_t *snapshots) {
    if (!adj) return fg::set_error(FG_INVALID, "fg_adjoint_timing: NULL handle");
    if (ms)
        for (int q = 0; q < 4; q++) ms[q] = adj->ms[q];
    if (snapshots) *snapshots = adj->n_snaps ? adj->n_snaps - 1 : 0;
    return FG_OK;
}

void fg_adjoint_free(fg_adjoint *adj) { delete adj; }

}  // extern "C"
