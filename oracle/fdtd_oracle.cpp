// FDTD ORACLE — test infrastructure only, never the product.
//
// A CPU restatement of the reference's FDTD path, used by the tests, by
// __graft_entry__.smoke() and by bench.py's cpu_baseline leg to check and time the HIP
// engine. Nothing in fieldgrid_amd/ links, loads or calls this file.
//
// It restates, in the reference's operation order:
//   crates/sims-field/src/fdtd/sim.rs   axis 29-131, time_step 133-137, psi_len 139-143,
//                                       Sim::try_new 352-554, thin_wire 556-635, wires 303-336,
//                                       pulse_shape/gaussian 646-657, psi_idx 659-665
//   crates/sims-field/src/fdtd/run.rs   CoefSets 61-100, port_loop 102-127, pack 205-376
//   crates/sims-field/src/fdtd/cpu.rs   State::step 267-307 and every routine it calls,
//                                       energy 417-421, run_cpu 424-482
//   crates/sims-field/src/fdtd/ntff.rs  patches 11-59
// Template parameter R is the field/coefficient type:
//   R = float  : the reference's FP32 path (f64 builder, f32 cast, f32 stepper, f64 probe sums)
//   R = double : the FP64 oracle (every value kept in f64)
// Extensions the reference lacks are marked "extension": point probes, per-cell Debye pole
// sets, and a planted-mistake switch (oracle_set_mutation) used only by the tests that show
// the pinning checks catch a wrong oracle.
//
// Pinning: tests/test_oracle_*.py and tests/test_antenna_dipole.py check this file against the
// reference's own unit tests (run.rs:775-858), its analytic cavity test (tests/fdtd.rs:189-236),
// its cancellation and wire-radius tests (tests/fdtd.rs:176-187, 271-286), its antenna bars
// (crates/sims-antenna/tests/fdtd.rs:45-55), and closed forms and exact discrete identities
// (Courant dt, CPML b/c, Yee dispersion, CPML reflection, Debye limits and mid-band poles
// against Debye::response, the port's Thevenin law, lumped C and L laws, the plane DFT against
// direct sums, the sheet impedance and face weights, dipole NTFF power balance), each shown
// to fail on a planted mistake (oracle_set_mutation 1-14).

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <set>
#include <string>
#include <tuple>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

#include "../include/fieldgrid_sim.h"

namespace {

constexpr double C0 = 299792458.0;
constexpr double EPS0 = 8.8541878128e-12;
constexpr double MU0 = 1.25663706212e-6;
constexpr double ETA0 = 376.730313668;
constexpr double PI = 3.14159265358979323846;

int g_mutation = 0;  // extension: 0 = faithful; >0 plants one known mistake (tests only)

size_t idx(const size_t n[3], size_t i, size_t j, size_t k) { return (i * n[1] + j) * n[2] + k; }

struct Err {
    std::string msg;
};

// sim.rs:12-27
template <class R>
struct Axis {
    size_t n = 0;
    std::vector<double> d, dd;
    std::vector<R> inv_d, inv_dd, raw_inv_d, raw_inv_dd, be, ce, bh, ch;
    std::vector<int> slot_e, slot_h;
};

// sim.rs:29-131
template <class R>
Axis<R> axis(const std::vector<double> &line, size_t pml, double dt) {
    size_t n = line.size();
    Axis<R> a;
    a.n = n;
    a.d.resize(n);
    a.dd.resize(n);
    for (size_t i = 0; i < n; i++) a.d[i] = (i + 1 < n) ? line[i + 1] - line[i] : line[i] - line[i - 1];
    for (size_t i = 0; i < n; i++) {
        if (i == 0)
            a.dd[i] = a.d[0];
        else if (i + 1 >= n)
            a.dd[i] = a.d[n - 2];
        else
            a.dd[i] = 0.5 * (line[i + 1] - line[i - 1]);
    }
    const double m = 3.0, kappa_max = 1.0, alpha_max = 0.05;
    std::vector<double> ke(n, 1.0), kh(n, 1.0);
    a.be.assign(n, R(0));
    a.ce.assign(n, R(0));
    a.bh.assign(n, R(0));
    a.ch.assign(n, R(0));
    a.slot_e.assign(n, -1);
    a.slot_h.assign(n, -1);
    if (pml > 0 && n > 2 * pml + 2) {
        double left = line[pml], right = line[n - 1 - pml];
        double depth_l = left - line[0], depth_r = line[n - 1] - right;
        auto coef = [&](double rho, double depth, double cell, double &kappa, double &b, double &c) {
            double r = std::min(std::max(rho / depth, 0.0), 1.0);
            double sigma_max = 0.8 * (m + 1.0) / (ETA0 * cell);
            double sigma = sigma_max * std::pow(r, m);
            kappa = 1.0 + (kappa_max - 1.0) * std::pow(r, m);
            double alpha = alpha_max * (1.0 - r) * 2.0 * PI * EPS0 * 1e9;
            b = std::exp(-(sigma / kappa + alpha) * dt / EPS0);
            c = sigma > 0.0 ? sigma * (b - 1.0) / (sigma * kappa + kappa * kappa * alpha) : 0.0;
            if (g_mutation == 3) std::swap(b, c);  // planted: CPML b and c transposed
        };
        for (size_t i = 0; i < n; i++) {
            double x = line[i];
            double rho, depth, cell;
            if (x < left) {
                rho = left - x; depth = depth_l; cell = a.d[0];
            } else if (x > right) {
                rho = x - right; depth = depth_r; cell = a.d[n - 2];
            } else {
                rho = 0.0; depth = 1.0; cell = 1.0;
            }
            if (rho > 0.0) {
                double k, b, c;
                coef(rho, depth, cell, k, b, c);
                ke[i] = k;
                a.be[i] = R(b);
                a.ce[i] = R(c);
            }
            if (i < pml)
                a.slot_e[i] = int(i);
            else if (i > n - 1 - pml)
                a.slot_e[i] = int(i + 2 * pml - n);
            if (i + 1 < n) {
                double xc = 0.5 * (line[i] + line[i + 1]);
                if (xc < left) {
                    rho = left - xc; depth = depth_l; cell = a.d[0];
                } else if (xc > right) {
                    rho = xc - right; depth = depth_r; cell = a.d[n - 2];
                } else {
                    rho = 0.0; depth = 1.0; cell = 1.0;
                }
                if (rho > 0.0) {
                    double k, b, c;
                    coef(rho, depth, cell, k, b, c);
                    kh[i] = k;
                    a.bh[i] = R(b);
                    a.ch[i] = R(c);
                }
                if (i < pml)
                    a.slot_h[i] = int(i);
                else if (i >= n - 1 - pml)
                    a.slot_h[i] = int(pml + i - (n - 1 - pml));
            }
        }
    }
    a.inv_d.resize(n);
    a.inv_dd.resize(n);
    a.raw_inv_d.resize(n);
    a.raw_inv_dd.resize(n);
    for (size_t i = 0; i < n; i++) {
        a.inv_d[i] = R(1.0 / (kh[i] * a.d[i]));
        a.inv_dd[i] = R(1.0 / (ke[i] * a.dd[i]));
        a.raw_inv_d[i] = R(1.0 / a.d[i]);
        a.raw_inv_dd[i] = R(1.0 / a.dd[i]);
    }
    return a;
}

// sim.rs:133-137
double time_step(const std::vector<double> lines[3]) {
    double s = 0.0;
    for (int a = 0; a < 3; a++) {
        double mn = INFINITY;
        for (size_t i = 0; i + 1 < lines[a].size(); i++) mn = std::min(mn, lines[a][i + 1] - lines[a][i]);
        s += 1.0 / (mn * mn);
    }
    return 0.99 / (C0 * std::sqrt(s));
}

// sim.rs:139-143
size_t psi_len(const size_t n[3], size_t a, size_t pml) {
    if (pml == 0) return 0;
    size_t m[3] = {n[0], n[1], n[2]};
    m[a] = 2 * pml;
    return m[0] * m[1] * m[2];
}

// sim.rs:659-665
size_t psi_idx(const size_t n[3], size_t a, size_t pml, const size_t p[3], size_t slot) {
    if (a == 0) return (slot * n[1] + p[1]) * n[2] + p[2];
    if (a == 1) return (p[0] * 2 * pml + slot) * n[2] + p[2];
    return (p[0] * n[1] + p[1]) * 2 * pml + slot;
}

// sim.rs:646-649
void pulse_shape(double fc, double &t0, double &tau) {
    tau = std::sqrt(std::log(10.0)) / (PI * fc);
    t0 = 4.0 * tau;
}

// Rust's `{:.3e}` for a positive finite f64: d.ddd e<exp>, no '+' and no zero padding.
std::string lower_exp3(double v) {
    char buf[64];
    std::snprintf(buf, sizeof buf, "%.3e", v);
    std::string s(buf);
    size_t e = s.find('e');
    std::string mant = s.substr(0, e);
    int ex = std::atoi(s.c_str() + e + 1);
    return mant + "e" + std::to_string(ex);
}

// sim.rs:303-313
double wire_factor(double d, double radius) {
    double lo = d * std::exp(-4.0), hi = 0.5 * d;
    if (radius >= lo && radius < hi) return 2.0 / std::log(d / radius);
    throw Err{"a thin wire's radius must be from a 55th to half of its cell (" + lower_exp3(lo) + " to " +
              lower_exp3(hi) + " m), not " + lower_exp3(radius) + " m"};
}

struct Edge {
    size_t comp;
    size_t at[3];
};

struct PortDef {
    double r;
    bool reference_above;
    std::vector<std::vector<Edge>> columns;
};

struct SheetEdge {
    size_t comp, id;
    double r;
    bool adaptive;
    size_t faces[2];
    double len, c;
};

template <class R>
struct WireH {
    size_t h;
    size_t e[4];
    R c[4];
};

// sim.rs:173-189
template <class R>
struct Sim {
    std::vector<double> line[3];
    size_t n[3];
    size_t nn;
    size_t pml;
    Axis<R> ax[3];
    double dt;
    std::vector<R> ca[3], cb[3];
    std::vector<PortDef> ports;
    std::vector<std::vector<std::tuple<size_t, size_t, R>>> port_src;
    std::vector<std::tuple<size_t, size_t, R, R>> inductors;
    std::vector<SheetEdge> sheets;
    std::vector<float> sheet_table;
    size_t sheet_states;
    std::vector<std::vector<std::pair<double, double>>> pole_sets;  // extension: >1 set
    std::vector<std::tuple<size_t, size_t, double, uint32_t>> debye_edges;
    std::vector<WireH<R>> wires;
};

Edge to_edge(const fg_edge &e) { return Edge{e.comp, {e.at[0], e.at[1], e.at[2]}}; }

// sim.rs:315-336
std::vector<double> wire_cells(const std::vector<double> lines[3], size_t pml, const size_t n[3],
                               const fg_sim_desc &d) {
    std::vector<double> out;
    for (uint32_t q = 0; q < d.n_wires; q++) {
        Edge e = to_edge(d.wires[q].edge);
        double radius = d.wires[q].radius;
        size_t w = e.comp;
        size_t p[3] = {e.at[0], e.at[1], e.at[2]};
        size_t end[3] = {p[0], p[1], p[2]};
        end[w] += 1;
        auto clear = [&](const size_t *x) {
            for (int a = 0; a < 3; a++)
                if (!(x[a] > pml && x[a] + 1 + pml < n[a])) return false;
            return true;
        };
        if (!clear(p) || !clear(end)) throw Err{"a thin wire must stay clear of the grid's absorbing layer"};
        for (size_t ax : {(w + 1) % 3, (w + 2) % 3}) {
            const std::vector<double> &line = lines[ax];
            for (const size_t *node : {(const size_t *)p, (const size_t *)end})
                for (size_t cell : {node[ax], node[ax] - 1}) out.push_back(wire_factor(line[cell + 1] - line[cell], radius));
        }
    }
    return out;
}

template <class R>
void thin_wire(Sim<R> &sim, const fg_sim_desc &d);

// sim.rs:352-554
template <class R>
Sim<R> build(const fg_sim_desc &d) {
    Sim<R> sim;
    for (int a = 0; a < 3; a++) {
        if (!d.line[a] || d.len[a] < 2) throw Err{"a grid line needs at least two nodes"};
        sim.line[a].assign(d.line[a], d.line[a] + d.len[a]);
        sim.n[a] = d.len[a];
    }
    const size_t *n = sim.n;
    size_t total = n[0] * n[1] * n[2];
    sim.nn = total;
    sim.pml = d.pml;
    double speedup = 1.0;
    for (double f : wire_cells(sim.line, d.pml, n, d)) speedup = std::max(speedup, std::max(f, 1.0 / f));
    double dt = time_step(sim.line) / std::sqrt(speedup);
    sim.dt = dt;
    for (int a = 0; a < 3; a++) sim.ax[a] = axis<R>(sim.line[a], d.pml, dt);
    const Axis<R> *ax = sim.ax;
    size_t cells[3] = {n[0] - 1, n[1] - 1, n[2] - 1};
    auto mat = [&](const long c[3], double out[4]) {
        long cc[3];
        for (int a = 0; a < 3; a++) cc[a] = std::min(std::max(c[a], 0L), long(n[a]) - 2);
        size_t id = (size_t(cc[0]) * cells[1] + size_t(cc[1])) * cells[2] + size_t(cc[2]);
        out[0] = d.eps ? double(d.eps[id]) : 1.0;
        out[1] = d.sigma ? double(d.sigma[id]) : 0.0;
        out[2] = d.inf ? double(d.inf[id]) : 1.0;
        out[3] = d.debye ? double(d.debye[id]) : 0.0;
    };
    for (int c = 0; c < 3; c++) {
        sim.ca[c].assign(total, R(0));
        sim.cb[c].assign(total, R(0));
    }
    std::vector<double> eps_edge[3], sig_edge[3], inf_edge[3], dd_edge[3];
    std::vector<uint32_t> set_edge[3];  // extension: pole set per edge
    for (int c = 0; c < 3; c++) {
        eps_edge[c].assign(total, 0.0);
        sig_edge[c].assign(total, 0.0);
        inf_edge[c].assign(total, 0.0);
        dd_edge[c].assign(total, 0.0);
        set_edge[c].assign(total, 0);
        size_t u = (c + 1) % 3, v = (c + 2) % 3;
        for (size_t i = 0; i < n[0]; i++)
            for (size_t j = 0; j < n[1]; j++)
                for (size_t k = 0; k < n[2]; k++) {
                    size_t p[3] = {i, j, k};
                    if (p[c] + 1 >= n[c] || p[u] == 0 || p[u] + 1 >= n[u] || p[v] == 0 || p[v] + 1 >= n[v]) continue;
                    double sum[4] = {0, 0, 0, 0};
                    double w_sum = 0.0;
                    double best = -1.0;
                    uint32_t set = 0;
                    const long offs[4][2] = {{-1, -1}, {0, -1}, {-1, 0}, {0, 0}};
                    for (auto &o : offs) {
                        long cell[3] = {long(i), long(j), long(k)};
                        cell[u] += o[0];
                        cell[v] += o[1];
                        double w = ax[u].d[size_t(cell[u])] * ax[v].d[size_t(cell[v])];
                        double m[4];
                        mat(cell, m);
                        for (int q = 0; q < 4; q++) sum[q] += m[q] * w;
                        w_sum += w;
                        if (d.debye_set) {  // extension: the set of the cell with most dd*w
                            long cc[3];
                            for (int a = 0; a < 3; a++) cc[a] = std::min(std::max(cell[a], 0L), long(n[a]) - 2);
                            size_t cid = (size_t(cc[0]) * cells[1] + size_t(cc[1])) * cells[2] + size_t(cc[2]);
                            if (m[3] * w > best) {
                                best = m[3] * w;
                                set = d.debye_set[cid];
                            }
                        }
                    }
                    size_t id = idx(n, i, j, k);
                    eps_edge[c][id] = EPS0 * sum[0] / w_sum;
                    sig_edge[c][id] = sum[1] / w_sum;
                    inf_edge[c][id] = sum[2] / w_sum;
                    dd_edge[c][id] = sum[3] / w_sum;
                    set_edge[c][id] = set;
                }
    }
    auto area = [&](const Edge &e) {
        size_t u = (e.comp + 1) % 3, v = (e.comp + 2) % 3;
        return ax[u].dd[e.at[u]] * ax[v].dd[e.at[v]];
    };
    auto length = [&](const Edge &e) { return ax[e.comp].d[e.at[e.comp]]; };
    auto eid = [&](const Edge &e) { return idx(n, e.at[0], e.at[1], e.at[2]); };
    std::vector<std::tuple<size_t, size_t, double>> series_r;
    for (uint32_t l = 0; l < d.n_lumped; l++) {
        const fg_lumped &lp = d.lumped[l];
        double count = double(std::max<uint32_t>(lp.n_edges, 1));
        for (uint32_t q = 0; q < lp.n_edges; q++) {
            Edge e = to_edge(lp.edges[q]);
            size_t id = eid(e);
            if (lp.kind == FG_CAPACITOR)
                eps_edge[e.comp][id] += g_mutation == 10 ? lp.value * count * area(e) / length(e)  // planted: A/l for l/A
                                                         : lp.value * count * length(e) / area(e);
            else if (lp.kind == FG_RESISTOR)
                series_r.emplace_back(e.comp, id, lp.value / count);
        }
    }
    for (int c = 0; c < 3; c++)
        for (size_t id = 0; id < total; id++) {
            double eps = eps_edge[c][id];
            if (eps == 0.0) continue;
            double x = sig_edge[c][id] * dt / (2.0 * eps);
            sim.ca[c][id] = R((1.0 - x) / (1.0 + x));
            sim.cb[c][id] = R(dt / eps / (1.0 + x));
        }
    if (d.pec)
        for (int c = 0; c < 3; c++)
            for (size_t id = 0; id < total; id++)
                if (d.pec[c * total + id]) {
                    sim.ca[c][id] = R(0);
                    sim.cb[c][id] = R(0);
                }
    std::set<std::pair<size_t, size_t>> excluded;
    for (uint32_t l = 0; l < d.n_lumped; l++)
        for (uint32_t q = 0; q < d.lumped[l].n_edges; q++) {
            Edge e = to_edge(d.lumped[l].edges[q]);
            excluded.insert({e.comp, eid(e)});
        }
    for (uint32_t p = 0; p < d.n_ports; p++) {
        const fg_port_def &pd = d.ports[p];
        PortDef def{pd.r, pd.reference_above != 0, {}};
        size_t o = 0;
        for (uint32_t c = 0; c < pd.n_columns; c++) {
            std::vector<Edge> col;
            for (uint32_t q = 0; q < pd.column_len[c]; q++) col.push_back(to_edge(pd.edges[o + q]));
            o += pd.column_len[c];
            for (auto &e : col) excluded.insert({e.comp, eid(e)});
            def.columns.push_back(col);
        }
        sim.ports.push_back(def);
    }
    for (uint32_t r = 0; r < d.n_resistive; r++) excluded.insert({d.resistive[r].comp, size_t(d.resistive[r].id)});
    // Debye pole sets: the reference has one global table (sim.rs:243-246)
    uint32_t nsets = d.n_pole_sets > 1 ? d.n_pole_sets : 1;
    sim.pole_sets.resize(nsets);
    for (uint32_t s = 0; s < nsets; s++) {
        uint32_t a = d.set_first ? d.set_first[s] : 0, b = d.set_first ? d.set_first[s + 1] : d.n_poles;
        for (uint32_t q = a; q < b; q++) sim.pole_sets[s].push_back({d.poles[2 * q], d.poles[2 * q + 1]});
    }
    std::vector<double> half_g(nsets);
    for (uint32_t s = 0; s < nsets; s++) {
        double acc = 0.0;
        for (auto [x, a] : sim.pole_sets[s]) {
            if (g_mutation == 9) x *= 2.0;  // planted: relaxation rate doubled (tau halved)
            acc += a * x * dt / (1.0 + 0.5 * x * dt);
        }
        half_g[s] = acc * 0.5;
    }
    for (int c = 0; c < 3; c++)
        for (size_t id = 0; id < total; id++) {
            double dd = dd_edge[c][id];
            if (dd <= 0.0 || sim.cb[c][id] == R(0) || excluded.count({size_t(c), id})) continue;
            uint32_t s = set_edge[c][id];
            double eps = EPS0 * (inf_edge[c][id] + dd * half_g[s]);
            sim.ca[c][id] = R(1.0);
            sim.cb[c][id] = R(dt / eps);
            sim.debye_edges.emplace_back(size_t(c), id, dd, s);
        }
    auto resistor = [&](const Edge &e, double r) -> double {
        size_t id = eid(e);
        double eps = eps_edge[e.comp][id];
        double x_sig = sig_edge[e.comp][id] * dt / (2.0 * eps);
        double beta = dt * length(e) / (2.0 * r * eps * area(e)) + x_sig;
        sim.ca[e.comp][id] = R((1.0 - beta) / (1.0 + beta));
        sim.cb[e.comp][id] = R(dt / (eps * (1.0 + beta)));
        return dt / (eps * (1.0 + beta) * r * area(e));
    };
    size_t nn = total;
    for (auto &[comp, id, r] : series_r) {
        Edge e{comp, {id / (n[1] * n[2]), (id / n[2]) % n[1], id % n[2]}};
        resistor(e, r);
    }
    for (uint32_t q = 0; q < d.n_resistive; q++) {
        const fg_resistive &s = d.resistive[q];
        size_t id = size_t(s.id);
        Edge e{s.comp, {id / (n[1] * n[2]), (id / n[2]) % n[1], id % n[2]}};
        sim.sheets.push_back(SheetEdge{s.comp, id, s.r, s.adaptive != 0, {size_t(s.faces[0]), size_t(s.faces[1])},
                                       length(e), double(sim.cb[s.comp][id]) / area(e)});
    }
    (void)nn;
    for (auto &p : sim.ports) {
        double cols = double(std::max<size_t>(p.columns.size(), 1));
        std::vector<std::tuple<size_t, size_t, R>> v;
        for (auto &col : p.columns) {
            double m = double(std::max<size_t>(col.size(), 1));
            double re = p.r * cols / m;
            for (auto &e : col) {
                double src = resistor(e, re);
                v.emplace_back(e.comp, eid(e), R(src / m));
            }
        }
        sim.port_src.push_back(v);
    }
    for (uint32_t l = 0; l < d.n_lumped; l++) {
        const fg_lumped &lp = d.lumped[l];
        if (lp.kind != FG_INDUCTOR) continue;
        double count = double(std::max<uint32_t>(lp.n_edges, 1));
        for (uint32_t q = 0; q < lp.n_edges; q++) {
            Edge e = to_edge(lp.edges[q]);
            size_t id = eid(e);
            double k_e = double(sim.cb[e.comp][id]) / area(e);
            double k_i = dt * length(e) / (g_mutation == 11 ? lp.value * count : lp.value / count);  // planted 11: L*count
            sim.inductors.emplace_back(e.comp, id, R(k_e), R(k_i));
        }
    }
    if (d.sheet_table && d.sheet_table_len)
        sim.sheet_table.assign(d.sheet_table, d.sheet_table + d.sheet_table_len);
    else
        sim.sheet_table.assign(4, 0.0f);
    sim.sheet_states = d.sheet_states;
    thin_wire(sim, d);
    return sim;
}

// sim.rs:556-635
template <class R>
void thin_wire(Sim<R> &sim, const fg_sim_desc &d) {
    const size_t *n = sim.n;
    size_t nn = sim.nn;
    double k_mu = sim.dt / MU0;
    auto key = [&](size_t c, const size_t p[3]) { return c * nn + idx(n, p[0], p[1], p[2]); };
    std::unordered_set<size_t> on_wire;
    for (uint32_t q = 0; q < d.n_wires; q++) {
        Edge e = to_edge(d.wires[q].edge);
        on_wire.insert(key(e.comp, e.at));
    }
    struct HS {
        size_t q[3];
        size_t c;
        double f;
    };
    std::map<size_t, HS> scaled_h;
    std::map<size_t, double> scaled_e;
    for (uint32_t qq = 0; qq < d.n_wires; qq++) {
        Edge e = to_edge(d.wires[qq].edge);
        double radius = d.wires[qq].radius;
        size_t w = e.comp;
        size_t p[3] = {e.at[0], e.at[1], e.at[2]};
        size_t end[3] = {p[0], p[1], p[2]};
        end[w] += 1;
        for (size_t axx : {(w + 1) % 3, (w + 2) % 3}) {
            size_t c = 3 - w - axx;
            size_t q1[3] = {p[0], p[1], p[2]};
            size_t q2[3] = {p[0], p[1], p[2]};
            q2[axx] -= 1;
            std::pair<size_t, const size_t *> hs[2] = {{p[axx], q1}, {p[axx] - 1, q2}};
            for (auto &[cell, q] : hs) {
                double f = wire_factor(sim.ax[axx].d[cell], radius);
                size_t k = key(c, q);
                if (!scaled_h.count(k)) scaled_h[k] = HS{{q[0], q[1], q[2]}, c, f};
            }
            for (const size_t *node : {(const size_t *)p, (const size_t *)end}) {
                size_t a1[3] = {node[0], node[1], node[2]};
                size_t a2[3] = {node[0], node[1], node[2]};
                a2[axx] -= 1;
                std::pair<size_t, const size_t *> es[2] = {{node[axx], a1}, {node[axx] - 1, a2}};
                for (auto &[cell, at] : es) {
                    size_t id = key(axx, at);
                    if (!on_wire.count(id)) {
                        double f = wire_factor(sim.ax[axx].d[cell], radius);
                        if (!scaled_e.count(id)) scaled_e[id] = f;
                    }
                }
            }
        }
    }
    for (auto &[id, f] : scaled_e) {
        size_t c = id / nn, at = id % nn;
        double ca = double(sim.ca[c][at]), cb = double(sim.cb[c][at]);
        if (cb == 0.0) continue;
        double a = (1.0 - ca) / (1.0 + ca);
        double scaled = a / f;
        sim.ca[c][at] = R((1.0 - scaled) / (1.0 + scaled));
        sim.cb[c][at] = R(cb * (1.0 + a) / (f * (1.0 + scaled)));
    }
    sim.wires.clear();
    for (auto &[h, hs] : scaled_h) {
        size_t c = hs.c, u = (c + 1) % 3, v = (c + 2) % 3;
        double s = (hs.f - 1.0) * k_mu;
        double du = double(sim.ax[u].inv_d[hs.q[u]]), dv = double(sim.ax[v].inv_d[hs.q[v]]);
        size_t qu[3] = {hs.q[0], hs.q[1], hs.q[2]};
        qu[u] += 1;
        size_t qv[3] = {hs.q[0], hs.q[1], hs.q[2]};
        qv[v] += 1;
        std::pair<size_t, double> all[4] = {
            {key(v, qu), -s * du}, {key(v, hs.q), s * du}, {key(u, qv), s * dv}, {key(u, hs.q), -s * dv}};
        WireH<R> wh;
        wh.h = h;
        for (int t = 0; t < 4; t++) {
            wh.e[t] = h;
            wh.c[t] = R(0);
        }
        int slot = 0;
        for (auto &[e, cc] : all)
            if (!on_wire.count(e)) {
                wh.e[slot] = e;
                wh.c[slot] = R(cc);
                slot++;
            }
        sim.wires.push_back(wh);
    }
}

// run.rs:102-127
template <class R>
std::vector<std::tuple<size_t, size_t, double>> port_loop(const Sim<R> &sim, const PortDef &p, size_t axs, size_t k) {
    const size_t *n = sim.n;
    size_t u = (axs + 1) % 3, v = (axs + 2) % 3;
    auto at = [&](size_t a, size_t b) {
        size_t q[3] = {0, 0, 0};
        q[u] = a;
        q[v] = b;
        q[axs] = k;
        return idx(n, q[0], q[1], q[2]);
    };
    std::map<std::pair<size_t, size_t>, int> turns;
    std::set<std::pair<size_t, size_t>> nodes;
    for (auto &c : p.columns) nodes.insert({c[0].at[u], c[0].at[v]});
    for (auto &[a, b] : nodes) {
        turns[{at(a, b), v}] += 1;
        turns[{at(a - 1, b), v}] -= 1;
        turns[{at(a, b), u}] -= 1;
        turns[{at(a, b - 1), u}] += 1;
    }
    std::vector<std::tuple<size_t, size_t, double>> out;
    for (auto &[key, t] : turns) {
        if (t == 0) continue;
        auto [id, comp] = key;
        size_t pc[3] = {id / (n[1] * n[2]), (id / n[2]) % n[1], id % n[2]};
        out.emplace_back(id, comp, double(t) * sim.ax[comp].dd[pc[comp]]);
    }
    return out;
}

// run.rs:61-100
struct CoefSets {
    std::vector<uint32_t> index;
    std::vector<float> sets;  // pairs
    bool wide;
};

CoefSets coef_sets(const Sim<float> &sim, size_t narrow) {
    size_t nn = sim.nn;
    std::map<std::pair<uint32_t, uint32_t>, uint32_t> lookup;
    CoefSets cs;
    cs.sets = {0.0f, 0.0f};
    lookup[{0, 0}] = 0;
    std::vector<uint32_t> flat;
    flat.reserve(3 * nn);
    std::pair<uint32_t, uint32_t> last_key{0, 0};
    uint32_t last_id = 0;
    for (int c = 0; c < 3; c++)
        for (size_t id = 0; id < nn; id++) {
            float a = sim.ca[c][id], b = sim.cb[c][id];
            std::pair<uint32_t, uint32_t> key{0, 0};
            if (b != 0.0f) {
                uint32_t ab, bb;
                std::memcpy(&ab, &a, 4);
                std::memcpy(&bb, &b, 4);
                key = {ab, bb};
            }
            if (key != last_key) {
                auto it = lookup.find(key);
                uint32_t sid;
                if (it == lookup.end()) {
                    cs.sets.push_back(a);
                    cs.sets.push_back(b);
                    sid = uint32_t(cs.sets.size() / 2 - 1);
                    lookup[key] = sid;
                } else {
                    sid = it->second;
                }
                last_key = key;
                last_id = sid;
            }
            flat.push_back(last_id);
        }
    cs.wide = cs.sets.size() / 2 > narrow;
    if (cs.wide) {
        cs.index = flat;
    } else {
        for (size_t q = 0; q < flat.size(); q += 2)
            cs.index.push_back(flat[q] | (q + 1 < flat.size() ? flat[q + 1] << 16 : 0u));
    }
    return cs;
}

// run.rs:178-203, with integer indices
template <class R>
struct Packed {
    size_t offsets[18];
    size_t psi_total;
    std::vector<size_t> source_id, source_comp;  // driven port
    std::vector<R> source_src;
    std::vector<R> inductor;  // 5 per: comp, id (as R only for layout), k_e, k_i, I
    std::vector<size_t> ind_f;
    std::vector<R> debye_table;  // per set: np x (ka, kb), padded
    size_t np;
    struct PE {
        size_t id;
        R w;
    };
    struct LT {
        size_t id, comp;
        R w;
    };
    struct PD {
        size_t comp, first, count, lf, lc;
    };
    std::vector<PE> port_edges;
    std::vector<LT> loops;
    std::vector<PD> probe_def;
    double t0, tau;
    size_t cap, ports, nf, plane_len, patches;
    std::vector<float> sheets;  // 16 per sheet (ids as bit patterns)
    size_t sheet_count;
};

template <class R>
Packed<R> pack(const Sim<R> &sim, size_t driven, const fg_pulse &pulse, const fg_extras &ex) {
    Packed<R> p;
    const size_t *n = sim.n;
    size_t nn = sim.nn;
    size_t total = 0;
    for (int e = 0; e < 2; e++)
        for (int c = 0; c < 3; c++)
            for (int a = 0; a < 3; a++) {
                p.offsets[e * 9 + c * 3 + a] = total;
                if (a != c) total += psi_len(n, a, sim.pml);
            }
    p.psi_total = total;
    if (driven >= sim.port_src.size() && !sim.port_src.empty()) throw Err{"no such port"};
    if (!sim.port_src.empty())
        for (auto &[comp, id, s] : sim.port_src[driven]) {
            p.source_id.push_back(id);
            p.source_comp.push_back(comp);
            p.source_src.push_back(s);
        }
    for (auto &[comp, id, ke, ki] : sim.inductors) {
        p.inductor.insert(p.inductor.end(), {R(0), R(0), ke, ki, R(0)});
        p.ind_f.push_back(comp * nn + id);
    }
    p.np = 0;
    for (auto &s : sim.pole_sets) p.np = std::max(p.np, s.size());
    for (auto &s : sim.pole_sets) {
        for (size_t q = 0; q < p.np; q++) {
            if (q < s.size()) {
                double x = s[q].first, a = s[q].second;
                if (g_mutation == 9) x *= 2.0;  // planted: relaxation rate doubled (tau halved)
                double dn = 1.0 + 0.5 * x * sim.dt;
                p.debye_table.push_back(R((1.0 - 0.5 * x * sim.dt) / dn));
                p.debye_table.push_back(R(EPS0 * a * x / dn));
            } else {
                p.debye_table.push_back(R(0));
                p.debye_table.push_back(R(0));
            }
        }
    }
    for (auto &pd : sim.ports) {
        size_t axs = pd.columns[0][0].comp;
        const auto &mid = pd.columns[pd.columns.size() / 2];
        size_t first = p.port_edges.size();
        double polarity = pd.reference_above ? -1.0 : 1.0;
        double weight = polarity / double(pd.columns.size());
        for (auto &col : pd.columns)
            for (auto &e : col)
                p.port_edges.push_back({idx(n, e.at[0], e.at[1], e.at[2]), R(sim.ax[axs].d[e.at[axs]] * weight)});
        size_t probed = p.port_edges.size() - first;
        size_t loop_first = p.loops.size();
        double height = 0.0;
        for (auto &e : mid) height += sim.ax[axs].d[e.at[axs]];
        for (auto &e : mid) {
            size_t k = e.at[axs];
            double share = polarity * sim.ax[axs].d[k] / height;
            for (auto &[id, comp, w] : port_loop(sim, pd, axs, k))
                p.loops.push_back({id, comp, R(w * share)});
        }
        p.probe_def.push_back({axs, first, probed, loop_first, p.loops.size() - loop_first});
    }
    pulse_shape(pulse.fc, p.t0, p.tau);
    p.cap = ex.max_steps;
    p.ports = sim.ports.size();
    p.nf = std::min<size_t>(ex.nf, 4);
    p.plane_len = ex.plane_k >= 0 ? n[0] * n[1] * p.nf * 10 : 0;
    p.patches = ex.n_patches;
    for (auto &s : sim.sheets) {
        size_t hc;
        bool ok;
        if (s.comp == 0) {
            hc = 1; ok = true;
        } else if (s.comp == 1) {
            hc = 0; ok = true;
        } else {
            hc = 0; ok = false;
        }
        if (!ok || s.id % n[2] == 0) continue;
        auto bits = [](size_t v) {
            uint32_t u = uint32_t(v);
            float f;
            std::memcpy(&f, &u, 4);
            return f;
        };
        float vals[16] = {bits(s.comp * nn + s.id), bits(hc * nn + s.id), bits(hc * nn + s.id - 1),
                          float(s.len), float(s.c), float(s.r), s.adaptive ? 1.0f : 0.0f, 0.0f, 0.0f, 0.0f,
                          0.5f, 0.5f, 0.0f, 0.0f, bits(s.faces[0]), bits(s.faces[1])};
        p.sheets.insert(p.sheets.end(), vals, vals + 16);
    }
    p.sheet_count = p.sheets.size() / 16;
    return p;
}

uint32_t fbits(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return u;
}

// cpu.rs:169-193: the face weights of one sheet record (f32 in the reference for both paths)
void sheet_weights(float *sh, float old, float ha, float hb) {
    sh[12] = old;
    float aa = sh[7] + ha * ha, bb = sh[8] + hb * hb, ab = sh[9] + ha * hb;
    sh[7] = aa;
    sh[8] = bb;
    sh[9] = ab;
    float sum = aa + bb;
    if (sum <= 1e-30f) return;
    float scale = 1.0f;
    if (sh[6] != 0.0f) {
        scale = sh[10] + sh[11];
        float tot = aa + bb - 2.0f * ab;
        if (tot > 1e-30f) scale = std::min(std::max(sum / tot, 0.5f), 4.0f);
    }
    if (g_mutation == 12) scale = 1.0f;  // planted: adaptive scale never applied
    sh[10] = scale * aa / sum;
    sh[11] = scale * bb / sum;
}

// cpu.rs:17-422
template <class R>
struct State {
    const Sim<R> &sim;
    Packed<R> p;
    const fg_extras &ex;
    std::vector<R> e, h, psi, inductor, current, debye_old, debye_state, series, plane, ntff, probes;
    std::vector<float> sheets;
    std::vector<R> omega;
    R w0;
    size_t plane_k;

    State(const Sim<R> &s, Packed<R> pk, const fg_extras &x, const fg_pulse &pulse)
        : sim(s), p(std::move(pk)), ex(x) {
        size_t nn = sim.nn;
        e.assign(3 * nn, R(0));
        h.assign(3 * nn, R(0));
        psi.assign(p.psi_total, R(0));
        inductor = p.inductor;
        sheets = p.sheets;
        current.assign(p.sheet_count * sim.sheet_states, R(0));
        debye_old.assign(sim.debye_edges.size(), R(0));
        debye_state.assign(sim.debye_edges.size() * p.np, R(0));
        series.assign(p.cap * p.ports * 2, R(0));
        plane.assign(p.plane_len, R(0));
        ntff.assign(p.patches * p.nf * 12, R(0));
        probes.assign(p.cap * ex.n_probes, R(0));
        for (size_t k = 0; k < p.nf; k++) omega.push_back(R(2.0 * PI * ex.freqs[k]));
        w0 = R(2.0 * PI * pulse.f0);
        plane_k = ex.plane_k >= 0 ? size_t(ex.plane_k) : 0;
    }

    size_t base(int ee, size_t c, size_t a) const { return p.offsets[ee * 9 + c * 3 + a]; }

    // cpu.rs:48-107
    void update_h() {
        const size_t *n = sim.n;
        size_t nn = sim.nn, pml = sim.pml;
        R k_mu = R(sim.dt / MU0);
        const Axis<R> *ax = sim.ax;
#pragma omp parallel for schedule(static)
        for (long ii = 0; ii < long(n[0]); ii++) {
            size_t i = size_t(ii);
            for (size_t j = 0; j < n[1]; j++)
                for (size_t k = 0; k < n[2]; k++) {
                    size_t pp[3] = {i, j, k};
                    size_t id = (i * n[1] + j) * n[2] + k;
                    for (size_t c = 0; c < 3; c++) {
                        size_t u = (c + 1) % 3, v = (c + 2) % 3;
                        if (pp[u] + 1 >= n[u] || pp[v] + 1 >= n[v]) continue;
                        size_t up[3] = {i, j, k}, vp[3] = {i, j, k};
                        up[u] += 1;
                        vp[v] += 1;
                        R ev = e[v * nn + idx(n, up[0], up[1], up[2])] - e[v * nn + id];
                        R eu = e[u * nn + idx(n, vp[0], vp[1], vp[2])] - e[u * nn + id];
                        size_t pu = pp[u], pv = pp[v];
                        R t1 = ev * ax[u].inv_d[pu];
                        R t2 = eu * ax[v].inv_d[pv];
                        if (g_mutation == 4) t2 = eu * ax[u].inv_d[pv];  // planted: wrong axis table
                        if (pml > 0 && g_mutation != 1) {  // planted 1: CPML psi dropped
                            int su = ax[u].slot_h[pu];
                            if (su >= 0) {
                                R &r = psi[base(1, c, u) + psi_idx(n, u, pml, pp, size_t(su))];
                                r = ax[u].bh[pu] * r + ax[u].ch[pu] * ev * ax[u].raw_inv_d[pu];
                                t1 += r;
                            }
                            int sv = ax[v].slot_h[pv];
                            if (sv >= 0) {
                                R &r = psi[base(1, c, v) + psi_idx(n, v, pml, pp, size_t(sv))];
                                r = ax[v].bh[pv] * r + ax[v].ch[pv] * eu * ax[v].raw_inv_d[pv];
                                t2 += r;
                            }
                        }
                        if (g_mutation == 2)
                            h[c * nn + id] += k_mu * (t1 - t2);  // planted: curl sign flipped
                        else
                            h[c * nn + id] -= k_mu * (t1 - t2);
                    }
                }
        }
    }

    // cpu.rs:109-167
    void update_e() {
        const size_t *n = sim.n;
        size_t nn = sim.nn, pml = sim.pml;
        const Axis<R> *ax = sim.ax;
#pragma omp parallel for schedule(static)
        for (long ii = 0; ii < long(n[0]); ii++) {
            size_t i = size_t(ii);
            for (size_t j = 0; j < n[1]; j++)
                for (size_t k = 0; k < n[2]; k++) {
                    size_t pp[3] = {i, j, k};
                    size_t id = (i * n[1] + j) * n[2] + k;
                    for (size_t c = 0; c < 3; c++) {
                        R cb = sim.cb[c][id];
                        if (cb == R(0)) continue;
                        size_t u = (c + 1) % 3, v = (c + 2) % 3;
                        size_t um[3] = {i, j, k}, vm[3] = {i, j, k};
                        um[u] -= 1;
                        vm[v] -= 1;
                        R hv = h[v * nn + id] - h[v * nn + idx(n, um[0], um[1], um[2])];
                        R hu = h[u * nn + id] - h[u * nn + idx(n, vm[0], vm[1], vm[2])];
                        size_t pu = pp[u], pv = pp[v];
                        R t1 = hv * ax[u].inv_dd[pu];
                        R t2 = hu * ax[v].inv_dd[pv];
                        if (pml > 0 && g_mutation != 1) {
                            int su = ax[u].slot_e[pu];
                            if (su >= 0) {
                                R &r = psi[base(0, c, u) + psi_idx(n, u, pml, pp, size_t(su))];
                                r = ax[u].be[pu] * r + ax[u].ce[pu] * hv * ax[u].raw_inv_dd[pu];
                                t1 += r;
                            }
                            int sv = ax[v].slot_e[pv];
                            if (sv >= 0) {
                                R &r = psi[base(0, c, v) + psi_idx(n, v, pml, pp, size_t(sv))];
                                r = ax[v].be[pv] * r + ax[v].ce[pv] * hu * ax[v].raw_inv_dd[pv];
                                t2 += r;
                            }
                        }
                        R &f = e[c * nn + id];
                        f = sim.ca[c][id] * f + cb * (t1 - t2);
                    }
                }
        }
    }

    static float bitsf(uint32_t u) {
        float f;
        std::memcpy(&f, &u, 4);
        return f;
    }

    void sheet_before(size_t s, float old, float ha, float hb) { sheet_weights(&sheets[s * 16], old, ha, hb); }

    // cpu.rs:195-237
    void sheet_after(size_t s) {
        float *sh = &sheets[s * 16];
        const std::vector<float> &br = sim.sheet_table;
        size_t f = fbits(sh[0]);
        float len = sh[3], c = sh[4], r = sh[5];
        float ma = sh[10], mb = sh[11], old = sh[12], i_old = sh[13];
        size_t nr = size_t(br[0]), nc = size_t(br[1]);
        size_t stride = 2 + nr + 2 * nc, faces = 4 + 2 * nr + 4 * nc;
        size_t fa = faces + fbits(sh[14]) * stride, fb = faces + fbits(sh[15]) * stride;
        size_t b = s * (nr + 2 * nc);
        R hist = R(ma * br[fa + 1] + mb * br[fb + 1]) * R(i_old);
        if (g_mutation == 13) hist = -hist;  // planted: sign of the i_old history term
        for (size_t k = 0; k < nr; k++) hist -= R(ma * br[fa + 2 + k] + mb * br[fb + 2 + k]) * current[b + k];
        for (size_t k = 0; k < nc; k++) {
            size_t o = 2 + nr + 2 * k;
            R cr = R(ma * br[fa + o] + mb * br[fb + o]);
            R ci = R(ma * br[fa + o + 1] + mb * br[fb + o + 1]);
            size_t u = b + nr + 2 * k;
            hist -= cr * current[u] - ci * current[u + 1];
        }
        R a = R(r * (ma * br[fa] + mb * br[fb]));
        R ibar = (R(0.5f * len) * (e[f] + R(old)) + R(r) * hist) / (a + R(0.5f * len * c));
        e[f] -= R(c) * ibar;
        sh[13] = float(R(2) * ibar - R(i_old));
        for (size_t k = 0; k < nr; k++) {
            size_t pq = 4 + 2 * k;
            current[b + k] = g_mutation == 14 ? R(br[pq + 1]) * current[b + k] + R(br[pq]) * ibar  // planted: p, q swapped
                                              : R(br[pq]) * current[b + k] + R(br[pq + 1]) * ibar;
        }
        for (size_t k = 0; k < nc; k++) {
            size_t pq = 4 + 2 * nr + 4 * k;
            size_t u = b + nr + 2 * k;
            R ur = current[u], ui = current[u + 1];
            current[u] = R(br[pq]) * ur - R(br[pq + 1]) * ui + R(br[pq + 2]) * ibar;
            current[u + 1] = R(br[pq]) * ui + R(br[pq + 1]) * ur + R(br[pq + 3]) * ibar;
        }
    }

    // cpu.rs:239-260
    void debye_after(size_t s) {
        size_t nn = sim.nn;
        auto &[c, id, dd64, set] = sim.debye_edges[s];
        size_t f = c * nn + id;
        R dd = R(dd64);
        R old = debye_old[s];
        R cb = sim.cb[c][id];
        const R *tab = &p.debye_table[size_t(set) * p.np * 2];
        size_t count = sim.debye_edges.size();
        R s0 = R(0);
        for (size_t k = 0; k < p.np; k++) s0 += R(0.5) * (R(1.0) + tab[2 * k]) * debye_state[k * count + s];
        R next = e[f] - cb * s0;
        e[f] = next;
        R de = next - old;
        for (size_t k = 0; k < p.np; k++) {
            size_t q = k * count + s;
            R kb = g_mutation == 5 ? tab[2 * k] : tab[2 * k + 1];  // planted: ka used for kb
            debye_state[q] = tab[2 * k] * debye_state[q] + dd * kb * de;
        }
    }

    // cpu.rs:262-265 (f32 in the reference; f64 in the FP64 oracle)
    R pulse(R t) const {
        R x = (t - R(p.t0)) / R(p.tau);
        return std::exp(-x * x) * std::cos(w0 * (t - R(p.t0)));
    }

    // cpu.rs:267-307
    void step(size_t nstep) {
        size_t nn = sim.nn;
        update_h();
        for (auto &w : sim.wires) {
            R s = R(0);
            for (int t = 0; t < 4; t++) s += w.c[t] * e[w.e[t]];
            h[w.h] += s;
        }
        for (size_t s = 0; s < p.sheet_count; s++) {
            const float *sh = &sheets[s * 16];
            size_t f = fbits(sh[0]);
            float ha = float(h[fbits(sh[1])]), hb = float(h[fbits(sh[2])]);
            sheet_before(s, float(e[f]), ha, hb);
        }
        for (size_t s = 0; s < sim.debye_edges.size(); s++) {
            auto &[c, id, dd, set] = sim.debye_edges[s];
            debye_old[s] = e[c * nn + id];
        }
        update_e();
        for (size_t s = 0; s < p.sheet_count; s++) sheet_after(s);
        for (size_t s = 0; s < sim.debye_edges.size(); s++) debye_after(s);
        R dt = R(sim.dt);
        R t = (R(double(nstep)) + R(0.5)) * dt;
        if (g_mutation == 6) t = (R(double(nstep)) + R(1.0)) * dt;  // planted: source half-step late
        R pl = pulse(t);
        for (size_t q = 0; q < p.source_id.size(); q++)
            e[p.source_comp[q] * nn + p.source_id[q]] -= p.source_src[q] * pl;
        for (size_t q = 0; q < p.ind_f.size(); q++) {
            R *iq = &inductor[q * 5];
            size_t f = p.ind_f[q];
            R x = e[f] - iq[2] * iq[4];
            e[f] = x;
            iq[4] += iq[3] * x;
        }
        probe(nstep);
        field_dft(nstep);
        far(nstep);
        point_probes(nstep);
    }

    // cpu.rs:309-334
    void probe(size_t nstep) {
        size_t nn = sim.nn, ports = p.ports;
        for (size_t k = 0; k < ports; k++) {
            auto &d = p.probe_def[k];
            double v = 0.0;
            for (size_t q = d.first; q < d.first + d.count; q++)
                v -= double(e[d.comp * nn + p.port_edges[q].id] * p.port_edges[q].w);
            double cur = 0.0;
            for (size_t q = d.lf; q < d.lf + d.lc; q++)
                cur += double(h[p.loops[q].comp * nn + p.loops[q].id] * p.loops[q].w);
            if (nstep < p.cap) {
                series[(nstep * ports + k) * 2] = R(v);
                series[(nstep * ports + k) * 2 + 1] = R(cur);
            }
        }
    }

    // cpu.rs:336-361
    void field_dft(size_t nstep) {
        if (p.plane_len == 0) return;
        const size_t *n = sim.n;
        size_t nn = sim.nn, nf = p.nf;
        R dt = R(sim.dt);
        R te = (R(double(nstep)) + R(1.0)) * dt;
        R th = (R(double(nstep)) + R(0.5)) * dt;
        if (g_mutation == 8) te = th;  // planted: plane-DFT E phase at the half step
#pragma omp parallel for schedule(static)
        for (long qq = 0; qq < long(n[0] * n[1]); qq++) {
            size_t q = size_t(qq);
            size_t i = q / n[1], j = q % n[1];
            size_t id = (i * n[1] + j) * n[2] + plane_k;
            R f[5] = {e[id], e[nn + id], e[2 * nn + id], h[id], h[nn + id]};
            R *out = &plane[q * nf * 10];
            for (size_t k = 0; k < nf; k++) {
                R w = omega[k];
                R ce0 = std::cos(w * te) * dt, ce1 = -std::sin(w * te) * dt;
                R ch0 = std::cos(w * th) * dt, ch1 = -std::sin(w * th) * dt;
                for (int c = 0; c < 5; c++) {
                    R p0 = c >= 3 ? ch0 : ce0, p1 = c >= 3 ? ch1 : ce1;
                    out[k * 10 + 2 * c] += f[c] * p0;
                    out[k * 10 + 2 * c + 1] += f[c] * p1;
                }
            }
        }
    }

    // cpu.rs:363-415
    void far(size_t nstep) {
        if (p.patches == 0 || p.nf == 0) return;
        size_t nn = sim.nn, nf = p.nf;
        R dt = R(sim.dt);
        R te = (R(double(nstep)) + R(1.0)) * dt;
        R th = (R(double(nstep)) + R(0.5)) * dt;
        const uint32_t *ix = ex.ntff;
#pragma omp parallel for schedule(static)
        for (long gg = 0; gg < long(p.patches); gg++) {
            size_t g = size_t(gg);
            size_t b = g * 13;
            uint32_t meta = ix[b];
            size_t axs = meta & 3;
            R side = (meta & 4) ? R(1) : R(-1);
            size_t u = (axs + 1) % 3, v = (axs + 2) % 3;
            auto at = [&](size_t k) { return size_t(ix[b + k]); };
            R ef[3] = {0, 0, 0}, hf[3] = {0, 0, 0};
            ef[u] = R(0.5) * (e[u * nn + at(1)] + e[u * nn + at(2)]);
            ef[v] = R(0.5) * (e[v * nn + at(3)] + e[v * nn + at(4)]);
            hf[u] = R(0.25) * (h[u * nn + at(5)] + h[u * nn + at(6)] + h[u * nn + at(7)] + h[u * nn + at(8)]);
            hf[v] = R(0.25) * (h[v * nn + at(9)] + h[v * nn + at(10)] + h[v * nn + at(11)] + h[v * nn + at(12)]);
            R nrm[3] = {0, 0, 0};
            nrm[axs] = side;
            R jv[3] = {nrm[1] * hf[2] - nrm[2] * hf[1], nrm[2] * hf[0] - nrm[0] * hf[2], nrm[0] * hf[1] - nrm[1] * hf[0]};
            R mv[3] = {ef[1] * nrm[2] - ef[2] * nrm[1], ef[2] * nrm[0] - ef[0] * nrm[2], ef[0] * nrm[1] - ef[1] * nrm[0]};
            R *out = &ntff[g * nf * 12];
            for (size_t k = 0; k < nf; k++) {
                R wf = omega[k];
                R pe0 = std::cos(-wf * te) * dt, pe1 = std::sin(-wf * te) * dt;
                R ph0 = std::cos(-wf * th) * dt, ph1 = std::sin(-wf * th) * dt;
                if (g_mutation == 7) pe1 = -pe1;  // planted: conjugated E phase
                size_t o = k * 12;
                for (int c = 0; c < 3; c++) {
                    out[o + 2 * c] += ph0 * jv[c];
                    out[o + 2 * c + 1] += ph1 * jv[c];
                    out[o + 6 + 2 * c] += pe0 * mv[c];
                    out[o + 6 + 2 * c + 1] += pe1 * mv[c];
                }
            }
        }
    }

    // extension: point probes record one field component per step
    void point_probes(size_t nstep) {
        if (nstep >= p.cap) return;
        const size_t *n = sim.n;
        size_t nn = sim.nn;
        for (uint32_t q = 0; q < ex.n_probes; q++) {
            const fg_point_probe &pr = ex.probes[q];
            size_t id = idx(n, pr.i, pr.j, pr.k);
            probes[nstep * ex.n_probes + q] = pr.field < 3 ? e[pr.field * nn + id] : h[(pr.field - 3) * nn + id];
        }
    }

    // cpu.rs:417-421
    double energy() const {
        double es = 0.0, hs = 0.0;
#pragma omp parallel for reduction(+ : es, hs)
        for (long q = 0; q < long(e.size()); q++) {
            double x = double(e[q]);
            double y = double(h[q]) * 376.73;
            es += x * x;
            hs += y * y;
        }
        return es + hs;
    }
};

struct Holder {
    Sim<float> sim;
    CoefSets cs;
    Packed<float> pk;
    std::vector<float> axes;
    std::vector<fg_port> ports;
    std::vector<fg_port_edge> port_edges;
    std::vector<fg_loop_term> loops;
    std::vector<fg_source> sources;
    std::vector<fg_inductor> inductors;
    std::vector<fg_debye_edge> debye;
    std::vector<float> debye_table;
    std::vector<fg_wire_h> wires;
};

thread_local std::string g_err;

}  // namespace

extern "C" {
typedef struct oracle_out {
    uint64_t steps;
    double decay_db;
    uint64_t progress_calls;
    double *series;  // caller-allocated, NULL = not wanted; sizes as fg_record
    double *plane;
    double *ntff;
    double *probes;
    double *e_out;
    double *h_out;
    double dt;
    double step_seconds;  // wall time of the stepping loop from step warm_steps on (no build, no pack)
    uint64_t warm_steps;  // in: steps stepped before the timed part of the loop starts
    double build_seconds;  // wall time of build + pack + state set-up
} oracle_out;

}

// run_cpu (cpu.rs:424-482) in precision 32 (the reference's path) or 64 (FP64 oracle).
// cancel_at > 0 makes the progress callback return false on that call (tests/fdtd.rs:176-187).
template <class R>
static int run_impl(const fg_sim_desc *desc, uint32_t driven, const fg_pulse *pulse, const fg_extras *ex,
                    uint64_t cancel_at, double dt_override, oracle_out *out) {
    const auto t_build = std::chrono::steady_clock::now();
    Sim<R> sim = build<R>(*desc);
    if (dt_override > 0) sim.dt = dt_override;  // the reference's pub `sim.dt` (sim.rs:176)
    out->dt = sim.dt;
    Packed<R> pk = pack(sim, driven, *pulse, *ex);
    State<R> st(sim, std::move(pk), *ex, *pulse);
    size_t chunk = ex->check_every == 0 ? 1000 : size_t(ex->check_every);
    size_t pulse_end = size_t(2.0 * st.p.t0 / sim.dt);
    size_t steps = 0;
    double peak = 0.0, decay = 0.0;
    uint64_t calls = 0;
    int status = FG_OK;
    auto t_start = std::chrono::steady_clock::now();
    out->build_seconds = std::chrono::duration<double>(t_start - t_build).count();
    while (steps < ex->max_steps) {
        size_t k = std::min<size_t>(chunk, ex->max_steps - steps);
        for (size_t s = 0; s < k; s++) {
            if (steps + s == out->warm_steps) t_start = std::chrono::steady_clock::now();
            st.step(steps + s);
        }
        steps += k;
        double e = st.energy();
        if (!std::isfinite(e)) {
            g_err = "the FDTD run went unstable";
            status = FG_UNSTABLE;
            break;
        }
        peak = std::max(peak, e);
        decay = (e > 0.0 && peak > 0.0) ? 10.0 * std::log10(peak / e) : 0.0;
        calls++;
        if (cancel_at && calls == cancel_at) {
            g_err = "cancelled";
            status = FG_CANCELLED;
            break;
        }
        if (steps > std::max(pulse_end, size_t(ex->min_steps)) && decay >= ex->decay_db) break;
    }
    out->step_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
    out->steps = steps;
    out->decay_db = decay;
    out->progress_calls = calls;
    auto copy = [](double *dst, const std::vector<R> &src, size_t count) {
        if (dst)
            for (size_t q = 0; q < count; q++) dst[q] = double(src[q]);
    };
    copy(out->series, st.series, std::min(st.series.size(), steps * st.p.ports * 2));
    copy(out->plane, st.plane, st.plane.size());
    copy(out->ntff, st.ntff, st.ntff.size());
    copy(out->probes, st.probes, std::min(st.probes.size(), steps * ex->n_probes));
    copy(out->e_out, st.e, st.e.size());
    copy(out->h_out, st.h, st.h.size());
    return status;
}

extern "C" {

const char *oracle_last_error(void) { return g_err.c_str(); }

void oracle_set_mutation(int m) { g_mutation = m; }

int oracle_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

int oracle_run(const fg_sim_desc *desc, uint32_t driven, const fg_pulse *pulse, const fg_extras *ex, int precision,
               uint64_t cancel_at, double dt_override, oracle_out *out) {
    try {
        if (precision == 64) return run_impl<double>(desc, driven, pulse, ex, cancel_at, dt_override, out);
        return run_impl<float>(desc, driven, pulse, ex, cancel_at, dt_override, out);
    } catch (const Err &e) {
        g_err = e.msg;
        return FG_INVALID;
    } catch (const std::exception &e) {
        g_err = e.what();
        return FG_INVALID;
    }
}

// The FP32 builder + pack restated as an fg_problem (run.rs:205-376 with integer indices),
// for comparing the product builder's tables field by field.
int oracle_pack(const fg_sim_desc *desc, uint64_t narrow, void **handle, fg_problem *out, double *dt) {
    try {
        Holder *hd = new Holder{build<float>(*desc)};
        Sim<float> &sim = hd->sim;
        *dt = sim.dt;
        hd->cs = coef_sets(sim, narrow ? narrow : 65536);
        fg_extras ex{};
        fg_pulse pl{1e9, 1e9};
        hd->pk = pack(sim, 0, pl, ex);
        for (int a = 0; a < 3; a++) {
            const Axis<float> &A = sim.ax[a];
            for (size_t i = 0; i < A.n; i++) {
                float row[10] = {A.inv_d[i], A.inv_dd[i], A.raw_inv_d[i], A.raw_inv_dd[i], A.be[i],
                                 A.ce[i], A.bh[i], A.ch[i], float(A.slot_e[i]), float(A.slot_h[i])};
                hd->axes.insert(hd->axes.end(), row, row + 10);
            }
        }
        for (size_t q = 0; q < hd->pk.probe_def.size(); q++) {
            auto &d = hd->pk.probe_def[q];
            fg_port fp{};
            fp.comp = uint32_t(d.comp);
            fp.first_edge = uint32_t(d.first);
            fp.n_edges = uint32_t(d.count);
            fp.first_loop = uint32_t(d.lf);
            fp.n_loops = uint32_t(d.lc);
            fp.first_source = uint32_t(hd->sources.size());
            fp.n_sources = uint32_t(sim.port_src[q].size());
            for (auto &[comp, id, s] : sim.port_src[q]) hd->sources.push_back(fg_source{id, uint32_t(comp), s});
            hd->ports.push_back(fp);
        }
        for (auto &pe : hd->pk.port_edges) hd->port_edges.push_back(fg_port_edge{pe.id, pe.w, 0});
        for (auto &lt : hd->pk.loops) hd->loops.push_back(fg_loop_term{lt.id, uint32_t(lt.comp), lt.w});
        for (auto &[comp, id, ke, ki] : sim.inductors) hd->inductors.push_back(fg_inductor{id, uint32_t(comp), ke, ki, 0.0f});
        for (auto &[c, id, dd, set] : sim.debye_edges)
            hd->debye.push_back(fg_debye_edge{c * sim.nn + id, float(dd), set});
        hd->debye_table = hd->pk.debye_table;
        for (auto &w : sim.wires) {
            fg_wire_h x{};
            x.h = w.h;
            for (int t = 0; t < 4; t++) {
                x.e[t] = w.e[t];
                x.c[t] = w.c[t];
            }
            hd->wires.push_back(x);
        }
        fg_problem pr{};
        for (int a = 0; a < 3; a++) pr.n[a] = uint32_t(sim.n[a]);
        pr.pml = uint32_t(sim.pml);
        pr.dt = sim.dt;
        pr.axes = hd->axes.data();
        pr.coef_wide = hd->cs.wide;
        pr.n_sets = uint32_t(hd->cs.sets.size() / 2);
        pr.coef_sets = hd->cs.sets.data();
        pr.coef_index = hd->cs.index.data();
        pr.n_ports = uint32_t(hd->ports.size());
        pr.n_sources = uint32_t(hd->sources.size());
        pr.n_port_edges = uint32_t(hd->port_edges.size());
        pr.n_loops = uint32_t(hd->loops.size());
        pr.ports = hd->ports.data();
        pr.port_edges = hd->port_edges.data();
        pr.loops = hd->loops.data();
        pr.sources = hd->sources.data();
        pr.n_inductors = uint32_t(hd->inductors.size());
        pr.inductors = hd->inductors.data();
        pr.n_debye = hd->debye.size();
        pr.debye_edges = hd->debye.data();
        pr.n_poles = uint32_t(hd->pk.np);
        pr.n_pole_sets = uint32_t(sim.pole_sets.size());
        pr.debye_table = hd->debye_table.data();
        pr.n_sheets = hd->pk.sheet_count;
        pr.sheets = hd->pk.sheets.data();
        pr.sheet_states = uint32_t(sim.sheet_states);
        pr.n_branches = uint32_t(sim.sheet_table.size());
        pr.branches = sim.sheet_table.data();
        pr.n_wires = hd->wires.size();
        pr.wires = hd->wires.data();
        *out = pr;
        *handle = hd;
        return FG_OK;
    } catch (const Err &e) {
        g_err = e.msg;
        return FG_INVALID;
    }
}

void oracle_pack_free(void *handle) { delete static_cast<Holder *>(handle); }

const float *oracle_pack_ca(void *handle, int c) { return static_cast<Holder *>(handle)->sim.ca[c].data(); }
const float *oracle_pack_cb(void *handle, int c) { return static_cast<Holder *>(handle)->sim.cb[c].data(); }

// ntff.rs:11-59
uint64_t oracle_patches(const double *lx, uint32_t nx, const double *ly, uint32_t ny, const double *lz, uint32_t nz,
                        uint32_t pml, uint32_t inset, uint32_t *gpu, double *pos, double *area) {
    const double *lines[3] = {lx, ly, lz};
    size_t n[3] = {nx, ny, nz};
    size_t pp = pml + inset;
    size_t lo[3] = {pp, pp, pp};
    size_t hi[3] = {n[0] - 1 - pp, n[1] - 1 - pp, n[2] - 1 - pp};
    uint64_t count = 0;
    for (size_t axs = 0; axs < 3; axs++) {
        size_t u = (axs + 1) % 3, v = (axs + 2) % 3;
        for (int sd = 0; sd < 2; sd++) {
            size_t plane = sd == 0 ? lo[axs] : hi[axs];
            for (size_t a = lo[u]; a < hi[u]; a++)
                for (size_t b = lo[v]; b < hi[v]; b++) {
                    if (pos) {
                        double *ps = &pos[count * 3];
                        ps[axs] = lines[axs][plane];
                        ps[u] = 0.5 * (lines[u][a] + lines[u][a + 1]);
                        ps[v] = 0.5 * (lines[v][b] + lines[v][b + 1]);
                    }
                    if (area) area[count] = (lines[u][a + 1] - lines[u][a]) * (lines[v][b + 1] - lines[v][b]);
                    if (gpu) {
                        uint32_t *g = &gpu[count * 13];
                        size_t q[3];
                        q[axs] = plane;
                        q[u] = a;
                        q[v] = b;
                        auto at = [&](std::initializer_list<std::pair<size_t, long>> moves) {
                            long r[3] = {long(q[0]), long(q[1]), long(q[2])};
                            for (auto &[ax, dd] : moves) r[ax] += dd;
                            return uint32_t(idx(n, size_t(r[0]), size_t(r[1]), size_t(r[2])));
                        };
                        g[0] = uint32_t(axs) | (sd == 1 ? 4u : 0u);
                        g[1] = at({});
                        g[2] = at({{v, 1}});
                        g[3] = at({});
                        g[4] = at({{u, 1}});
                        g[5] = at({});
                        g[6] = at({{axs, -1}});
                        g[7] = at({{u, 1}});
                        g[8] = at({{axs, -1}, {u, 1}});
                        g[9] = at({});
                        g[10] = at({{axs, -1}});
                        g[11] = at({{v, 1}});
                        g[12] = at({{axs, -1}, {v, 1}});
                    }
                    count++;
                }
        }
    }
    return count;
}

// The reference's port_loop (run.rs:102-127) for its own unit test (run.rs:775-823):
// writes (id, comp, w) triples and returns the term count.
uint64_t oracle_port_loop(const fg_sim_desc *desc, uint32_t port, uint32_t axs, uint32_t k, uint64_t *ids,
                          uint32_t *comps, double *ws) {
    try {
        Sim<float> sim = build<float>(*desc);
        auto terms = port_loop(sim, sim.ports[port], axs, k);
        for (size_t q = 0; q < terms.size(); q++) {
            auto &[id, comp, w] = terms[q];
            if (ids) ids[q] = id;
            if (comps) comps[q] = uint32_t(comp);
            if (ws) ws[q] = w;
        }
        return terms.size();
    } catch (const Err &e) {
        g_err = e.msg;
        return 0;
    }
}

// The CPML axis tables of sim.rs:29-131 in f64 (kappa, b, c before the f32 cast), for the
// closed-form checks: writes n rows of (d, dd, be, ce, bh, ch) as built at precision 64.
void oracle_axis64(const double *line, uint32_t n, uint32_t pml, double dt, double *rows) {
    std::vector<double> l(line, line + n);
    Axis<double> a = axis<double>(l, pml, dt);
    for (size_t i = 0; i < n; i++) {
        double r[8] = {a.d[i], a.dd[i], a.be[i], a.ce[i], a.bh[i], a.ch[i], double(a.slot_e[i]), double(a.slot_h[i])};
        std::memcpy(&rows[i * 8], r, sizeof r);
    }
}

// The sheet face weights of cpu.rs:169-193 over a sequence of (ha, hb) face fields, for the
// closed-form checks of the weights: a record starting with weights (ma0, mb0) and zero sums,
// adaptive or not; writes the weights after each step.
void oracle_sheet_weights(int adaptive, float ma0, float mb0, const float *ha, const float *hb, uint64_t n,
                          float *ma, float *mb) {
    float sh[16] = {0};
    sh[6] = adaptive ? 1.0f : 0.0f;
    sh[10] = ma0;
    sh[11] = mb0;
    for (uint64_t q = 0; q < n; q++) {
        sheet_weights(sh, 0.0f, ha[q], hb[q]);
        ma[q] = sh[10];
        mb[q] = sh[11];
    }
}

// Sim::try_new's error text only (tests/fdtd.rs:271-286).
int oracle_check(const fg_sim_desc *desc) {
    try {
        build<float>(*desc);
        return FG_OK;
    } catch (const Err &e) {
        g_err = e.msg;
        return FG_INVALID;
    }
}

}  // extern "C"

// ---- discrete adjoint of the FP64 step (an extension: the reference has no gradient path) --
//
// The observations are point probes of E after every step, s_n = E_c(node) (the point_probes
// extension above). Given dJ/ds_n the adjoint returns dJ/d eps_r for every cell of the design box
// [lo, hi). The step is linear in (E, H, psi) and its coefficients depend on eps only through
// (ca, cb) of the edges whose material average (sim.rs:380-410) reads a design cell:
//   dJ/dca_e = sum_n lamE_e^{n+1} E_e^n,  dJ/dcb_e = sum_n lamE_e^{n+1} curl_e^n,
// with lam^{n} = M^T lam^{n+1} + dJ/dx^n. M^T is written here by transposing each statement of
// the step (cpu.rs:267-307) in reverse order, in scatter form, statement for statement: the
// inductors (cpu.rs:298-303; their currents are state), the Debye ADE (cpu.rs:239-260; the pole
// states are state, E^n enters through E_old), update_e (cpu.rs:109-167), the thin wires
// (cpu.rs:270-273) and update_h (cpu.rs:48-107); the HIP engine uses a gather form derived
// separately. Resistive sheets are refused: their face weights follow the accumulated H^2 of
// the run (cpu.rs:169-193), so the step is not linear in the fields.
namespace {

struct DesignEdge {
    size_t c, id;
    double eps, sig;  // edge averages (F/m, S/m) before the f32 cast
    size_t cell[4];   // design-box cell index or SIZE_MAX for a cell outside the box
    double w[4];      // area weight / sum of weights
};

std::vector<DesignEdge> design_edges(const fg_sim_desc &d, const Sim<double> &sim, const uint32_t lo[3],
                                     const uint32_t hi[3]) {
    const size_t *n = sim.n;
    size_t cells[3] = {n[0] - 1, n[1] - 1, n[2] - 1};
    for (int a = 0; a < 3; a++)
        if (lo[a] >= hi[a] || hi[a] > cells[a]) throw Err{"the design box lies outside the grid"};
    size_t box[3] = {hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]};
    std::set<std::pair<size_t, size_t>> seen, excluded;
    for (uint32_t l = 0; l < d.n_lumped; l++)
        for (uint32_t q = 0; q < d.lumped[l].n_edges; q++) {
            Edge e = to_edge(d.lumped[l].edges[q]);
            excluded.insert({e.comp, idx(n, e.at[0], e.at[1], e.at[2])});
        }
    for (uint32_t p = 0; p < d.n_ports; p++) {
        size_t o = 0;
        for (uint32_t c = 0; c < d.ports[p].n_columns; c++)
            for (uint32_t q = 0; q < d.ports[p].column_len[c]; q++, o++) {
                Edge e = to_edge(d.ports[p].edges[o]);
                excluded.insert({e.comp, idx(n, e.at[0], e.at[1], e.at[2])});
            }
    }
    for (uint32_t r = 0; r < d.n_resistive; r++) excluded.insert({d.resistive[r].comp, size_t(d.resistive[r].id)});
    std::set<std::pair<size_t, size_t>> debye;
    for (auto &[c, id, dd, set] : sim.debye_edges) debye.insert({c, id});
    for (uint32_t q = 0; q < d.n_wires; q++) {  // the edges and H nodes around a wire are rescaled
        const fg_edge &e = d.wires[q].edge;
        bool near = true;
        for (int a = 0; a < 3; a++) near = near && long(e.at[a]) + 2 >= long(lo[a]) && long(e.at[a]) <= long(hi[a]) + 1;
        if (near) throw Err{"the design box touches a thin wire"};
    }
    std::vector<DesignEdge> out;
    for (size_t ci = lo[0]; ci < hi[0]; ci++)
        for (size_t cj = lo[1]; cj < hi[1]; cj++)
            for (size_t ck = lo[2]; ck < hi[2]; ck++)
                for (size_t c = 0; c < 3; c++) {
                    size_t u = (c + 1) % 3, v = (c + 2) % 3;
                    for (size_t du = 0; du < 2; du++)
                        for (size_t dv = 0; dv < 2; dv++) {
                            size_t p[3] = {ci, cj, ck};
                            p[u] += du;
                            p[v] += dv;
                            size_t id = idx(n, p[0], p[1], p[2]);
                            if (!seen.insert({c, id}).second) continue;
                            if (sim.cb[c][id] == 0.0 || debye.count({c, id})) continue;  // eps does not enter
                            if (excluded.count({c, id}))
                                throw Err{"the design box touches a port, lumped or resistive edge"};
                            DesignEdge e{c, id, 0, 0, {}, {}};
                            double sum_e = 0, sum_s = 0, w_sum = 0;
                            const long offs[4][2] = {{-1, -1}, {0, -1}, {-1, 0}, {0, 0}};
                            for (int q = 0; q < 4; q++) {
                                long cell[3] = {long(p[0]), long(p[1]), long(p[2])};
                                cell[u] += offs[q][0];
                                cell[v] += offs[q][1];
                                double w = sim.ax[u].d[size_t(cell[u])] * sim.ax[v].d[size_t(cell[v])];
                                size_t cid = (size_t(cell[0]) * cells[1] + size_t(cell[1])) * cells[2] + size_t(cell[2]);
                                sum_e += (d.eps ? double(d.eps[cid]) : 1.0) * w;
                                sum_s += (d.sigma ? double(d.sigma[cid]) : 0.0) * w;
                                w_sum += w;
                                bool in = true;
                                for (int a = 0; a < 3; a++) in = in && size_t(cell[a]) >= lo[a] && size_t(cell[a]) < hi[a];
                                e.cell[q] = in ? ((size_t(cell[0]) - lo[0]) * box[1] + (size_t(cell[1]) - lo[1])) * box[2] +
                                                     (size_t(cell[2]) - lo[2])
                                               : SIZE_MAX;
                                e.w[q] = w;
                            }
                            for (int q = 0; q < 4; q++) e.w[q] /= w_sum;
                            e.eps = EPS0 * sum_e / w_sum;
                            e.sig = sum_s / w_sum;
                            out.push_back(e);
                        }
                }
    return out;
}

int design_gradient(const fg_sim_desc *desc, uint32_t driven, const fg_pulse *pulse, uint64_t steps, uint32_t n_obs,
                    const fg_point_probe *obs, const uint32_t lo[3], const uint32_t hi[3], const double *dseries,
                    double *grad, double *series_out) {
    Sim<double> sim = build<double>(*desc);
    if (!sim.sheets.empty()) throw Err{"resistive sheets make the step nonlinear in the fields"};
    for (uint32_t q = 0; q < n_obs; q++)
        if (obs[q].field > 2) throw Err{"observations are E components"};
    fg_extras ex{};
    ex.max_steps = steps;
    ex.decay_db = INFINITY;
    ex.plane_k = -1;
    ex.n_probes = n_obs;
    ex.probes = obs;
    Packed<double> pk = pack(sim, driven, *pulse, ex);
    State<double> st(sim, std::move(pk), ex, *pulse);
    const std::vector<DesignEdge> de = design_edges(*desc, sim, lo, hi);
    const size_t ne = de.size(), nn = sim.nn;
    const size_t *n = sim.n;
    std::vector<double> tape_e(steps * ne), tape_c(steps * ne);
    for (size_t s = 0; s < steps; s++) {
        for (size_t q = 0; q < ne; q++) tape_e[s * ne + q] = st.e[de[q].c * nn + de[q].id];
        st.step(s);
        for (size_t q = 0; q < ne; q++) {  // E^{n+1} = ca E^n + cb curl^n on these edges
            size_t f = de[q].c * nn + de[q].id;
            tape_c[s * ne + q] = (st.e[f] - sim.ca[de[q].c][de[q].id] * tape_e[s * ne + q]) / sim.cb[de[q].c][de[q].id];
        }
    }
    if (series_out)
        for (size_t q = 0; q < steps * n_obs; q++) series_out[q] = st.probes[q];
    std::vector<double> le(3 * nn, 0.0), lh(3 * nn, 0.0), lp(st.psi.size(), 0.0);
    std::vector<double> gca(ne, 0.0), gcb(ne, 0.0);
    const size_t nd = sim.debye_edges.size(), np = st.p.np, nind = st.p.ind_f.size();
    std::vector<double> lj(np * nd, 0.0), dold(nd, 0.0), li(nind, 0.0);
    const Axis<double> *ax = sim.ax;
    const size_t pml = sim.pml;
    const double k_mu = sim.dt / MU0;
    for (size_t s = steps; s-- > 0;) {
        for (uint32_t q = 0; q < n_obs; q++)  // s_n = E after step n
            le[obs[q].field * nn + idx(n, obs[q].i, obs[q].j, obs[q].k)] += dseries[s * n_obs + q];
        for (size_t q = 0; q < ne; q++) {
            double l = le[de[q].c * nn + de[q].id];
            gca[q] += l * tape_e[s * ne + q];
            gcb[q] += l * tape_c[s * ne + q];
        }
        // transpose of the inductors, last first (cpu.rs:298-303):
        //   x = E - k_e I; E = x; I += k_i x
        for (size_t q = nind; q-- > 0;) {
            const size_t f = st.p.ind_f[q];
            const double k_e = st.inductor[q * 5 + 2], k_i = st.inductor[q * 5 + 3];
            le[f] += k_i * li[q];
            li[q] -= (g_mutation == 17 ? -k_e : k_e) * le[f];  // planted 17: k_e sign
        }
        // transpose of the Debye ADE (cpu.rs:239-260): next = E' - cb sum_k (1+ka_k)/2 J_k,
        // J_k <- ka_k J_k + dd kb_k (next - E_old), E_old = E^n
        for (size_t d = 0; d < nd; d++) {
            auto &[c, id, dd, set] = sim.debye_edges[d];
            const size_t f = c * nn + id;
            const double *tab = &st.p.debye_table[size_t(set) * np * 2];
            const double cb = sim.cb[c][id];
            double mu = le[f], back = 0.0;
            for (size_t k = 0; k < np; k++) back += dd * tab[2 * k + 1] * lj[k * nd + d];
            if (g_mutation == 15) back = 0.0;  // planted: pole states left out of the adjoint
            mu += back;      // adjoint of next, through E^{n+1} and the pole states
            dold[d] = -back;  // adjoint of E_old = E^n
            for (size_t k = 0; k < np; k++) lj[k * nd + d] = tab[2 * k] * lj[k * nd + d] - cb * 0.5 * (1.0 + tab[2 * k]) * mu;
            le[f] = mu;  // adjoint of E' (the E update's output)
        }
        // transpose of update_e: E_c = ca E_c + cb (t1 - t2), t1 = hv inv_dd_u + psi_u', ...
        for (size_t i = 0; i < n[0]; i++)
            for (size_t j = 0; j < n[1]; j++)
                for (size_t k = 0; k < n[2]; k++) {
                    size_t pp[3] = {i, j, k};
                    size_t id = idx(n, i, j, k);
                    for (size_t c = 0; c < 3; c++) {
                        double cb = sim.cb[c][id];
                        if (cb == 0.0) continue;
                        size_t u = (c + 1) % 3, v = (c + 2) % 3;
                        double a = cb * le[c * nn + id];
                        le[c * nn + id] *= sim.ca[c][id];
                        size_t pu = pp[u], pv = pp[v];
                        double hv = a * ax[u].inv_dd[pu], hu = -a * ax[v].inv_dd[pv];
                        if (pml > 0) {
                            if (ax[u].slot_e[pu] >= 0) {
                                double &r = lp[st.base(0, c, u) + psi_idx(n, u, pml, pp, size_t(ax[u].slot_e[pu]))];
                                double tot = r + a;  // adjoint of psi_u' (later steps + this t1)
                                hv += ax[u].ce[pu] * ax[u].raw_inv_dd[pu] * tot;
                                r = ax[u].be[pu] * tot;
                            }
                            if (ax[v].slot_e[pv] >= 0) {
                                double &r = lp[st.base(0, c, v) + psi_idx(n, v, pml, pp, size_t(ax[v].slot_e[pv]))];
                                double tot = r - a;
                                hu += ax[v].ce[pv] * ax[v].raw_inv_dd[pv] * tot;
                                r = ax[v].be[pv] * tot;
                            }
                        }
                        size_t um[3] = {i, j, k}, vm[3] = {i, j, k};
                        um[u] -= 1;
                        vm[v] -= 1;
                        lh[v * nn + id] += hv;  // hv = H_v(p) - H_v(p - u)
                        lh[v * nn + idx(n, um[0], um[1], um[2])] -= hv;
                        lh[u * nn + id] += hu;  // hu = H_u(p) - H_u(p - v)
                        lh[u * nn + idx(n, vm[0], vm[1], vm[2])] -= hu;
                    }
                }
        for (size_t d = 0; d < nd; d++) {
            auto &[c, id, dd, set] = sim.debye_edges[d];
            le[c * nn + id] += dold[d];
        }
        // transpose of the thin wires (cpu.rs:270-273): H[h] += sum_t c_t E^n[e_t]
        if (g_mutation != 16)  // planted 16: wires left out of the adjoint
            for (auto &w : sim.wires)
                for (int t = 0; t < 4; t++) le[w.e[t]] += w.c[t] * lh[w.h];
        // transpose of update_h: H_c -= k_mu (t1 - t2), t1 = ev inv_d_u + psi_u', ...
        for (size_t i = 0; i < n[0]; i++)
            for (size_t j = 0; j < n[1]; j++)
                for (size_t k = 0; k < n[2]; k++) {
                    size_t pp[3] = {i, j, k};
                    size_t id = idx(n, i, j, k);
                    for (size_t c = 0; c < 3; c++) {
                        size_t u = (c + 1) % 3, v = (c + 2) % 3;
                        if (pp[u] + 1 >= n[u] || pp[v] + 1 >= n[v]) continue;
                        double b = -k_mu * lh[c * nn + id];
                        size_t pu = pp[u], pv = pp[v];
                        double ev = b * ax[u].inv_d[pu], eu = -b * ax[v].inv_d[pv];
                        if (pml > 0) {
                            if (ax[u].slot_h[pu] >= 0) {
                                double &r = lp[st.base(1, c, u) + psi_idx(n, u, pml, pp, size_t(ax[u].slot_h[pu]))];
                                double tot = r + b;
                                ev += ax[u].ch[pu] * ax[u].raw_inv_d[pu] * tot;
                                r = ax[u].bh[pu] * tot;
                            }
                            if (ax[v].slot_h[pv] >= 0) {
                                double &r = lp[st.base(1, c, v) + psi_idx(n, v, pml, pp, size_t(ax[v].slot_h[pv]))];
                                double tot = r - b;
                                eu += ax[v].ch[pv] * ax[v].raw_inv_d[pv] * tot;
                                r = ax[v].bh[pv] * tot;
                            }
                        }
                        size_t up[3] = {i, j, k}, vp[3] = {i, j, k};
                        up[u] += 1;
                        vp[v] += 1;
                        le[v * nn + idx(n, up[0], up[1], up[2])] += ev;  // ev = E_v(p + u) - E_v(p)
                        le[v * nn + id] -= ev;
                        le[u * nn + idx(n, vp[0], vp[1], vp[2])] += eu;  // eu = E_u(p + v) - E_u(p)
                        le[u * nn + id] -= eu;
                    }
                }
    }
    // chain rule: ca = (1-x)/(1+x), cb = dt/(eps (1+x)), x = sig dt / (2 eps) (sim.rs:431-441)
    size_t box = size_t(hi[0] - lo[0]) * (hi[1] - lo[1]) * (hi[2] - lo[2]);
    for (size_t q = 0; q < box; q++) grad[q] = 0.0;
    const double dt = sim.dt;
    for (size_t q = 0; q < ne; q++) {
        const double eps = de[q].eps, x = de[q].sig * dt / (2.0 * eps);
        const double dca = 2.0 * x / (eps * (1.0 + x) * (1.0 + x)), dcb = -dt / (eps * eps * (1.0 + x) * (1.0 + x));
        const double g = gca[q] * dca + gcb[q] * dcb;  // dJ / d eps_edge (F/m)
        for (int t = 0; t < 4; t++)
            if (de[q].cell[t] != SIZE_MAX) grad[de[q].cell[t]] += g * EPS0 * de[q].w[t];
    }
    return FG_OK;
}

}  // namespace

extern "C" int oracle_design_gradient(const fg_sim_desc *desc, uint32_t driven, const fg_pulse *pulse, uint64_t steps,
                                      uint32_t n_obs, const fg_point_probe *obs, const uint32_t lo[3], const uint32_t hi[3],
                                      const double *dseries, double *grad, double *series_out) {
    try {
        return design_gradient(desc, driven, pulse, steps, n_obs, obs, lo, hi, dseries, grad, series_out);
    } catch (const Err &e) {
        g_err = e.msg;
        return FG_INVALID;
    }
}
