// Host-side problem builder: Sim::try_new + pack + patches, the C++ mirror of
//   crates/sims-field/src/fdtd/sim.rs:29-665, run.rs:61-127 and 205-376, ntff.rs:11-59.
//
// Same arithmetic as the reference (f64 build, f32 tables, the same operation order, so the
// tables are bit-identical), different shape: the per-edge material average is computed on
// the fly inside a multi-threaded sweep instead of through four 3*nn f64 temporaries
// (sim.rs:371-374, about 22 GB at 226 M nodes), and every index is an integer.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <set>
#include <string>
#include <thread>
#include <tuple>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "../../include/fieldgrid_sim.h"
#include "common.h"

namespace fg {

constexpr double C0 = 299792458.0;
constexpr double EPS0 = 8.8541878128e-12;
constexpr double MU0 = 1.25663706212e-6;
constexpr double ETA0 = 376.730313668;
constexpr double PI = 3.14159265358979323846;

struct BuildError {
    std::string msg;
};

template <class F>
static void parallel_for(size_t count, F &&body) {
    unsigned workers = std::max(1u, std::min(64u, std::thread::hardware_concurrency()));
    if (count < 2 || workers == 1) {
        for (size_t q = 0; q < count; q++) body(q, 0u);
        return;
    }
    workers = unsigned(std::min<size_t>(workers, count));
    std::vector<std::thread> pool;
    for (unsigned w = 0; w < workers; w++)
        pool.emplace_back([&, w] {
            size_t a = count * w / workers, b = count * (w + 1) / workers;
            for (size_t q = a; q < b; q++) body(q, w);
        });
    for (auto &t : pool) t.join();
}

static unsigned worker_count(size_t count) {
    unsigned workers = std::max(1u, std::min(64u, std::thread::hardware_concurrency()));
    return unsigned(std::min<size_t>(workers, std::max<size_t>(count, 1)));
}

// One axis of CPML and spacing tables (sim.rs:12-131).
struct AxisT {
    size_t n = 0;
    std::vector<double> d, dd;
    std::vector<float> inv_d, inv_dd, raw_inv_d, raw_inv_dd, be, ce, bh, ch;
    std::vector<int> slot_e, slot_h;
};

static void cpml(double rho, double depth, double cell, double dt, double &kappa, double &b, double &c) {
    const double m = 3.0, kappa_max = 1.0, alpha_max = 0.05;
    double r = std::min(std::max(rho / depth, 0.0), 1.0);
    double sigma_max = 0.8 * (m + 1.0) / (ETA0 * cell);
    double sigma = sigma_max * std::pow(r, m);
    kappa = 1.0 + (kappa_max - 1.0) * std::pow(r, m);
    double alpha = alpha_max * (1.0 - r) * 2.0 * PI * EPS0 * 1e9;
    b = std::exp(-(sigma / kappa + alpha) * dt / EPS0);
    c = sigma > 0.0 ? sigma * (b - 1.0) / (sigma * kappa + kappa * kappa * alpha) : 0.0;
}

static AxisT make_axis(const std::vector<double> &line, size_t pml, double dt) {
    AxisT a;
    size_t n = a.n = line.size();
    a.d.resize(n);
    a.dd.resize(n);
    for (size_t i = 0; i < n; i++) a.d[i] = i + 1 < n ? line[i + 1] - line[i] : line[i] - line[i - 1];
    for (size_t i = 0; i < n; i++)
        a.dd[i] = i == 0 ? a.d[0] : (i + 1 >= n ? a.d[n - 2] : 0.5 * (line[i + 1] - line[i - 1]));
    std::vector<double> ke(n, 1.0), kh(n, 1.0);
    a.be.assign(n, 0.f);
    a.ce.assign(n, 0.f);
    a.bh.assign(n, 0.f);
    a.ch.assign(n, 0.f);
    a.slot_e.assign(n, -1);
    a.slot_h.assign(n, -1);
    if (pml > 0 && n > 2 * pml + 2) {
        const double left = line[pml], right = line[n - 1 - pml];
        const double depth_l = left - line[0], depth_r = line[n - 1] - right;
        auto place = [&](double x, double &rho, double &depth, double &cell) {
            if (x < left) {
                rho = left - x, depth = depth_l, cell = a.d[0];
            } else if (x > right) {
                rho = x - right, depth = depth_r, cell = a.d[n - 2];
            } else {
                rho = 0.0, depth = 1.0, cell = 1.0;
            }
        };
        for (size_t i = 0; i < n; i++) {
            double rho, depth, cell, k, b, c;
            place(line[i], rho, depth, cell);
            if (rho > 0.0) {
                cpml(rho, depth, cell, dt, k, b, c);
                ke[i] = k, a.be[i] = float(b), a.ce[i] = float(c);
            }
            if (i < pml)
                a.slot_e[i] = int(i);
            else if (i > n - 1 - pml)
                a.slot_e[i] = int(i + 2 * pml - n);
            if (i + 1 < n) {
                place(0.5 * (line[i] + line[i + 1]), rho, depth, cell);
                if (rho > 0.0) {
                    cpml(rho, depth, cell, dt, k, b, c);
                    kh[i] = k, a.bh[i] = float(b), a.ch[i] = float(c);
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
        a.inv_d[i] = float(1.0 / (kh[i] * a.d[i]));
        a.inv_dd[i] = float(1.0 / (ke[i] * a.dd[i]));
        a.raw_inv_d[i] = float(1.0 / a.d[i]);
        a.raw_inv_dd[i] = float(1.0 / a.dd[i]);
    }
    return a;
}

// Rust's `{:.3e}` for a positive f64.
static std::string exp3(double v) {
    char buf[64];
    std::snprintf(buf, sizeof buf, "%.3e", v);
    std::string s(buf);
    size_t e = s.find('e');
    return s.substr(0, e) + "e" + std::to_string(std::atoi(s.c_str() + e + 1));
}

// sim.rs:303-313
static double wire_factor(double d, double radius) {
    const double lo = d * std::exp(-4.0), hi = 0.5 * d;
    if (radius >= lo && radius < hi) return 2.0 / std::log(d / radius);
    throw BuildError{"a thin wire's radius must be from a 55th to half of its cell (" + exp3(lo) + " to " + exp3(hi) +
                     " m), not " + exp3(radius) + " m"};
}

struct Sheet {
    uint32_t comp;
    uint64_t id;
    double r;
    bool adaptive;
    uint64_t faces[2];
    double len, c;
};

}  // namespace fg

struct fg_sim {
    std::vector<double> line[3];
    uint64_t n[3];
    uint64_t nn;
    uint32_t pml;
    double dt;
    fg::AxisT ax[3];
    std::vector<float> ca, cb;  // flat c*nn+id
    // ports (sim.rs:165-171) and their sources (sim.rs:510-523)
    struct Port {
        double r;
        bool above;
        std::vector<std::vector<fg_edge>> columns;
    };
    std::vector<Port> ports;
    std::vector<std::vector<fg_source>> port_src;
    std::vector<fg_inductor> inductors;
    std::vector<fg::Sheet> sheets;
    std::vector<float> sheet_table;
    uint32_t sheet_states;
    std::vector<std::vector<std::pair<double, double>>> pole_sets;
    std::vector<fg_debye_edge> debye;
    std::vector<fg_wire_h> wires;
    // pack() products, owned here
    std::vector<float> axes;
    std::vector<float> coef_sets;
    std::vector<uint32_t> coef_index;
    std::vector<fg_port> pk_ports;
    std::vector<fg_port_edge> pk_edges;
    std::vector<fg_loop_term> pk_loops;
    std::vector<fg_source> pk_sources;
    std::vector<float> pk_debye_table;
    std::vector<float> pk_sheets;
};

namespace fg {

static inline uint64_t node(const uint64_t n[3], uint64_t i, uint64_t j, uint64_t k) { return (i * n[1] + j) * n[2] + k; }

struct EdgeMat {
    double eps, sig, inf, dd;
    uint32_t set;
};

// The 4-cell area-weighted edge average of sim.rs:375-410 for one edge.
struct Averager {
    const fg_sim_desc &d;
    const AxisT *ax;
    uint64_t n[3];
    uint64_t cells[3];
    EdgeMat at(uint32_t c, uint64_t i, uint64_t j, uint64_t k) const {
        const uint32_t u = (c + 1) % 3, v = (c + 2) % 3;
        const uint64_t p[3] = {i, j, k};
        EdgeMat m{0, 0, 0, 0, 0};
        if (p[c] + 1 >= n[c] || p[u] == 0 || p[u] + 1 >= n[u] || p[v] == 0 || p[v] + 1 >= n[v]) return m;
        double sum[4] = {0, 0, 0, 0}, w_sum = 0.0, best = -1.0;
        static const int offs[4][2] = {{-1, -1}, {0, -1}, {-1, 0}, {0, 0}};
        for (int q = 0; q < 4; q++) {
            int64_t cell[3] = {int64_t(i), int64_t(j), int64_t(k)};
            cell[u] += offs[q][0];
            cell[v] += offs[q][1];
            const double w = ax[u].d[size_t(cell[u])] * ax[v].d[size_t(cell[v])];
            for (int a = 0; a < 3; a++) cell[a] = std::min<int64_t>(std::max<int64_t>(cell[a], 0), int64_t(n[a]) - 2);
            const uint64_t cid = (uint64_t(cell[0]) * cells[1] + uint64_t(cell[1])) * cells[2] + uint64_t(cell[2]);
            const double vals[4] = {d.eps ? double(d.eps[cid]) : 1.0, d.sigma ? double(d.sigma[cid]) : 0.0,
                                    d.inf ? double(d.inf[cid]) : 1.0, d.debye ? double(d.debye[cid]) : 0.0};
            for (int r = 0; r < 4; r++) sum[r] += vals[r] * w;
            w_sum += w;
            if (d.debye_set && vals[3] * w > best) {
                best = vals[3] * w;
                m.set = d.debye_set[cid];
            }
        }
        m.eps = EPS0 * sum[0] / w_sum;
        m.sig = sum[1] / w_sum;
        m.inf = sum[2] / w_sum;
        m.dd = sum[3] / w_sum;
        return m;
    }
};

static void build(fg_sim &s, const fg_sim_desc &d) {
    for (int a = 0; a < 3; a++) {
        if (!d.line[a] || d.len[a] < 2) throw BuildError{"a grid line needs at least two nodes"};
        s.line[a].assign(d.line[a], d.line[a] + d.len[a]);
        s.n[a] = d.len[a];
    }
    const uint64_t *n = s.n;
    const uint64_t nn = s.nn = n[0] * n[1] * n[2];
    s.pml = d.pml;
    // wire speed-up of dt (sim.rs:362-364, wire_cells 315-336)
    double speedup = 1.0;
    for (uint32_t q = 0; q < d.n_wires; q++) {
        const fg_edge &e = d.wires[q].edge;
        const uint32_t w = e.comp;
        uint64_t p[3] = {e.at[0], e.at[1], e.at[2]}, end[3] = {p[0], p[1], p[2]};
        end[w] += 1;
        for (const uint64_t *x : {(const uint64_t *)p, (const uint64_t *)end})
            for (int a = 0; a < 3; a++)
                if (!(x[a] > d.pml && x[a] + 1 + d.pml < n[a]))
                    throw BuildError{"a thin wire must stay clear of the grid's absorbing layer"};
        for (uint32_t axx : {(w + 1) % 3, (w + 2) % 3})
            for (const uint64_t *x : {(const uint64_t *)p, (const uint64_t *)end})
                for (uint64_t cell : {x[axx], x[axx] - 1}) {
                    const double f = wire_factor(s.line[axx][cell + 1] - s.line[axx][cell], d.wires[q].radius);
                    speedup = std::max(speedup, std::max(f, 1.0 / f));
                }
    }
    // time_step (sim.rs:133-137)
    double sum = 0.0;
    for (int a = 0; a < 3; a++) {
        double mn = INFINITY;
        for (size_t i = 0; i + 1 < s.line[a].size(); i++) mn = std::min(mn, s.line[a][i + 1] - s.line[a][i]);
        sum += 1.0 / (mn * mn);
    }
    const double dt = s.dt = 0.99 / (C0 * std::sqrt(sum)) / std::sqrt(speedup);
    for (int a = 0; a < 3; a++) s.ax[a] = make_axis(s.line[a], d.pml, dt);
    const AxisT *ax = s.ax;
    Averager avg{d, ax, {n[0], n[1], n[2]}, {n[0] - 1, n[1] - 1, n[2] - 1}};
    auto area = [&](const fg_edge &e) {
        const uint32_t u = (e.comp + 1) % 3, v = (e.comp + 2) % 3;
        return ax[u].dd[e.at[u]] * ax[v].dd[e.at[v]];
    };
    auto length = [&](const fg_edge &e) { return ax[e.comp].d[e.at[e.comp]]; };
    auto flat = [&](const fg_edge &e) { return e.comp * nn + node(n, e.at[0], e.at[1], e.at[2]); };
    for (uint32_t l = 0; l < d.n_lumped; l++)
        for (uint32_t q = 0; q < d.lumped[l].n_edges; q++)
            for (int a = 0; a < 3; a++)
                if (d.lumped[l].edges[q].comp > 2 || d.lumped[l].edges[q].at[a] >= n[a])
                    throw BuildError{"a lumped element edge lies outside the grid"};
    // capacitor additions to eps_edge, applied in the reference's order (sim.rs:418-430)
    std::unordered_map<uint64_t, std::vector<double>> cap_add;
    std::vector<std::pair<uint64_t, double>> series_r;
    for (uint32_t l = 0; l < d.n_lumped; l++) {
        const fg_lumped &lp = d.lumped[l];
        const double count = double(std::max<uint32_t>(lp.n_edges, 1));
        for (uint32_t q = 0; q < lp.n_edges; q++) {
            const fg_edge &e = lp.edges[q];
            if (lp.kind == FG_CAPACITOR)
                cap_add[flat(e)].push_back(lp.value * count * length(e) / area(e));
            else if (lp.kind == FG_RESISTOR)
                series_r.push_back({flat(e), lp.value / count});
        }
    }
    auto edge_mat = [&](uint64_t f) {
        const uint32_t c = uint32_t(f / nn);
        const uint64_t id = f % nn;
        EdgeMat m = avg.at(c, id / (n[1] * n[2]), (id / n[2]) % n[1], id % n[2]);
        auto it = cap_add.find(f);
        if (it != cap_add.end())
            for (double x : it->second) m.eps += x;
        return m;
    };
    // excluded edges (sim.rs:455-468)
    std::unordered_set<uint64_t> excluded;
    for (uint32_t l = 0; l < d.n_lumped; l++)
        for (uint32_t q = 0; q < d.lumped[l].n_edges; q++) excluded.insert(flat(d.lumped[l].edges[q]));
    for (uint32_t p = 0; p < d.n_ports; p++) {
        const fg_port_def &pd = d.ports[p];
        fg_sim::Port port{pd.r, pd.reference_above != 0, {}};
        size_t o = 0;
        for (uint32_t c = 0; c < pd.n_columns; c++) {
            std::vector<fg_edge> col(pd.edges + o, pd.edges + o + pd.column_len[c]);
            o += pd.column_len[c];
            for (auto &e : col) {
                for (int a = 0; a < 3; a++)
                    if (e.comp > 2 || e.at[a] >= n[a]) throw BuildError{"a port edge lies outside the grid"};
                excluded.insert(flat(e));
            }
            port.columns.push_back(std::move(col));
        }
        if (port.columns.empty() || port.columns[0].empty()) throw BuildError{"a port needs at least one edge"};
        s.ports.push_back(std::move(port));
    }
    for (uint32_t r = 0; r < d.n_resistive; r++) excluded.insert(d.resistive[r].comp * nn + d.resistive[r].id);
    // Debye pole sets (sim.rs:243-246; more than one set is an extension)
    const uint32_t nsets = d.n_pole_sets > 1 ? d.n_pole_sets : 1;
    s.pole_sets.assign(nsets, {});
    std::vector<double> half_g(nsets, 0.0);
    for (uint32_t q = 0; q < nsets; q++) {
        const uint32_t a = d.set_first ? d.set_first[q] : 0, b = d.set_first ? d.set_first[q + 1] : d.n_poles;
        double acc = 0.0;
        for (uint32_t p = a; p < b; p++) {
            const double x = d.poles[2 * p], w = d.poles[2 * p + 1];
            s.pole_sets[q].push_back({x, w});
            acc += w * x * dt / (1.0 + 0.5 * x * dt);
        }
        half_g[q] = acc * 0.5;
    }
    // the sweep: ca/cb, PEC, Debye edges (sim.rs:375-410, 431-454, 469-483)
    s.ca.assign(3 * nn, 0.f);
    s.cb.assign(3 * nn, 0.f);
    const uint64_t planes = n[0];
    std::vector<std::vector<fg_debye_edge>> debye_by_plane(3 * planes);
    parallel_for(3 * planes, [&](size_t q, unsigned) {
        const uint32_t c = uint32_t(q / planes);
        const uint64_t i = q % planes;
        auto &out = debye_by_plane[q];
        for (uint64_t j = 0; j < n[1]; j++)
            for (uint64_t k = 0; k < n[2]; k++) {
                const uint64_t id = node(n, i, j, k), f = c * nn + id;
                EdgeMat m = avg.at(c, i, j, k);
                if (!cap_add.empty()) {
                    auto it = cap_add.find(f);
                    if (it != cap_add.end())
                        for (double x : it->second) m.eps += x;
                }
                float a = 0.f, b = 0.f;
                if (m.eps != 0.0) {
                    const double x = m.sig * dt / (2.0 * m.eps);
                    a = float((1.0 - x) / (1.0 + x));
                    b = float(dt / m.eps / (1.0 + x));
                }
                if (d.pec && d.pec[f]) a = b = 0.f;
                if (m.dd > 0.0 && b != 0.f && !excluded.count(f)) {
                    const double eps = EPS0 * (m.inf + m.dd * half_g[m.set]);
                    a = 1.f;
                    b = float(dt / eps);
                    out.push_back(fg_debye_edge{f, float(m.dd), m.set});
                }
                s.ca[f] = a;
                s.cb[f] = b;
            }
    });
    for (auto &v : debye_by_plane) s.debye.insert(s.debye.end(), v.begin(), v.end());
    // lumped resistors and ports (sim.rs:484-523)
    auto resistor = [&](uint64_t f, const fg_edge &e, double r) {
        const EdgeMat m = edge_mat(f);
        const double x_sig = m.sig * dt / (2.0 * m.eps);
        const double beta = dt * length(e) / (2.0 * r * m.eps * area(e)) + x_sig;
        s.ca[f] = float((1.0 - beta) / (1.0 + beta));
        s.cb[f] = float(dt / (m.eps * (1.0 + beta)));
        return dt / (m.eps * (1.0 + beta) * r * area(e));
    };
    auto edge_of = [&](uint64_t f) {
        const uint64_t id = f % nn;
        return fg_edge{uint32_t(f / nn), {uint32_t(id / (n[1] * n[2])), uint32_t((id / n[2]) % n[1]), uint32_t(id % n[2])}};
    };
    for (auto &[f, r] : series_r) resistor(f, edge_of(f), r);
    for (uint32_t q = 0; q < d.n_resistive; q++) {
        const fg_resistive &rs = d.resistive[q];
        const uint64_t f = rs.comp * nn + rs.id;
        const fg_edge e = edge_of(f);
        s.sheets.push_back(Sheet{rs.comp, rs.id, rs.r, rs.adaptive != 0, {rs.faces[0], rs.faces[1]}, length(e),
                                 double(s.cb[f]) / area(e)});
    }
    for (auto &p : s.ports) {
        const double cols = double(std::max<size_t>(p.columns.size(), 1));
        std::vector<fg_source> v;
        for (auto &col : p.columns) {
            const double m = double(std::max<size_t>(col.size(), 1));
            const double re = p.r * cols / m;
            for (auto &e : col) {
                const double src = resistor(flat(e), e, re);
                v.push_back(fg_source{node(n, e.at[0], e.at[1], e.at[2]), e.comp, float(src / m)});
            }
        }
        s.port_src.push_back(std::move(v));
    }
    // inductors (sim.rs:524-534)
    for (uint32_t l = 0; l < d.n_lumped; l++) {
        const fg_lumped &lp = d.lumped[l];
        if (lp.kind != FG_INDUCTOR) continue;
        const double count = double(std::max<uint32_t>(lp.n_edges, 1));
        for (uint32_t q = 0; q < lp.n_edges; q++) {
            const fg_edge &e = lp.edges[q];
            const uint64_t f = flat(e);
            const double k_e = double(s.cb[f]) / area(e);
            const double k_i = dt * length(e) / (lp.value / count);
            s.inductors.push_back(fg_inductor{f % nn, e.comp, float(k_e), float(k_i), 0.f});
        }
    }
    if (d.sheet_table && d.sheet_table_len)
        s.sheet_table.assign(d.sheet_table, d.sheet_table + d.sheet_table_len);
    else
        s.sheet_table.assign(4, 0.f);
    s.sheet_states = d.sheet_states;
    // thin wires (sim.rs:556-635)
    if (d.n_wires) {
        const double k_mu = dt / MU0;
        auto key = [&](uint32_t c, const uint64_t p[3]) { return c * nn + node(n, p[0], p[1], p[2]); };
        std::unordered_set<uint64_t> on_wire;
        for (uint32_t q = 0; q < d.n_wires; q++) on_wire.insert(flat(d.wires[q].edge));
        struct HS {
            uint64_t q[3];
            uint32_t c;
            double f;
        };
        std::map<uint64_t, HS> scaled_h;
        std::map<uint64_t, double> scaled_e;
        for (uint32_t qq = 0; qq < d.n_wires; qq++) {
            const fg_edge &e = d.wires[qq].edge;
            const double radius = d.wires[qq].radius;
            const uint32_t w = e.comp;
            const uint64_t p[3] = {e.at[0], e.at[1], e.at[2]};
            uint64_t end[3] = {p[0], p[1], p[2]};
            end[w] += 1;
            for (uint32_t axx : {(w + 1) % 3, (w + 2) % 3}) {
                const uint32_t c = 3 - w - axx;
                for (int side = 0; side < 2; side++) {
                    uint64_t q[3] = {p[0], p[1], p[2]};
                    q[axx] -= side;
                    const double f = wire_factor(ax[axx].d[p[axx] - side], radius);
                    scaled_h.emplace(key(c, q), HS{{q[0], q[1], q[2]}, c, f});
                }
                for (const uint64_t *nd : {p, (const uint64_t *)end})
                    for (int side = 0; side < 2; side++) {
                        uint64_t at[3] = {nd[0], nd[1], nd[2]};
                        at[axx] -= side;
                        const uint64_t id = key(axx, at);
                        if (!on_wire.count(id)) {
                            const double f = wire_factor(ax[axx].d[nd[axx] - side], radius);
                            scaled_e.emplace(id, f);
                        }
                    }
            }
        }
        for (auto &[f, fac] : scaled_e) {
            const double ca = double(s.ca[f]), cb = double(s.cb[f]);
            if (cb == 0.0) continue;
            const double a = (1.0 - ca) / (1.0 + ca);
            const double scaled = a / fac;
            s.ca[f] = float((1.0 - scaled) / (1.0 + scaled));
            s.cb[f] = float(cb * (1.0 + a) / (fac * (1.0 + scaled)));
        }
        for (auto &[h, hs] : scaled_h) {
            const uint32_t c = hs.c, u = (c + 1) % 3, v = (c + 2) % 3;
            const double sc = (hs.f - 1.0) * k_mu;
            const double du = double(ax[u].inv_d[hs.q[u]]), dv = double(ax[v].inv_d[hs.q[v]]);
            uint64_t qu[3] = {hs.q[0], hs.q[1], hs.q[2]}, qv[3] = {hs.q[0], hs.q[1], hs.q[2]};
            qu[u] += 1;
            qv[v] += 1;
            const std::pair<uint64_t, double> all[4] = {
                {key(v, qu), -sc * du}, {key(v, hs.q), sc * du}, {key(u, qv), sc * dv}, {key(u, hs.q), -sc * dv}};
            fg_wire_h wh{};
            wh.h = h;
            for (int t = 0; t < 4; t++) wh.e[t] = h, wh.c[t] = 0.f;
            int slot = 0;
            for (auto &[eid, cc] : all)
                if (!on_wire.count(eid)) wh.e[slot] = eid, wh.c[slot] = float(cc), slot++;
            s.wires.push_back(wh);
        }
    }
}

// run.rs:102-127: the current loop around a port's columns at layer k along `axis`.
static std::vector<fg_loop_term> port_loop(const fg_sim &s, const fg_sim::Port &p, uint32_t axs, uint32_t k,
                                           double share) {
    const uint64_t *n = s.n;
    const uint32_t u = (axs + 1) % 3, v = (axs + 2) % 3;
    auto at = [&](uint64_t a, uint64_t b) {
        uint64_t q[3];
        q[u] = a, q[v] = b, q[axs] = k;
        return node(n, q[0], q[1], q[2]);
    };
    std::map<std::pair<uint64_t, uint32_t>, int> turns;
    std::set<std::pair<uint64_t, uint64_t>> nodes;
    for (auto &c : p.columns) nodes.insert({c[0].at[u], c[0].at[v]});
    for (auto &[a, b] : nodes) {
        turns[{at(a, b), v}] += 1;
        turns[{at(a - 1, b), v}] -= 1;
        turns[{at(a, b), u}] -= 1;
        turns[{at(a, b - 1), u}] += 1;
    }
    std::vector<fg_loop_term> out;
    for (auto &[key, t] : turns) {
        if (t == 0) continue;
        const uint64_t id = key.first;
        const uint32_t comp = key.second;
        const uint64_t pc[3] = {id / (n[1] * n[2]), (id / n[2]) % n[1], id % n[2]};
        const double w = double(t) * s.ax[comp].dd[pc[comp]];
        out.push_back(fg_loop_term{id, comp, float(w * share)});
    }
    return out;
}

// CoefSets (run.rs:61-100): set ids in order of first appearance over flat c*nn+id.
static void coef_sets(fg_sim &s, uint64_t narrow) {
    const uint64_t total = 3 * s.nn;
    // each worker dedupes a contiguous range in order; merging the local tables range by
    // range keeps the global first-appearance order of the sequential scan
    const unsigned workers = worker_count(total / 4096 + 1);
    std::vector<std::vector<uint64_t>> local_keys(workers);
    std::vector<std::vector<uint32_t>> local_flat(workers);
    auto key_of = [&](uint64_t f) -> uint64_t {
        if (s.cb[f] == 0.f) return 0;
        uint32_t a, b;
        std::memcpy(&a, &s.ca[f], 4);
        std::memcpy(&b, &s.cb[f], 4);
        return (uint64_t(a) << 32) | b;
    };
    parallel_for(workers, [&](size_t w, unsigned) {
        const uint64_t a = total * w / workers, b = total * (w + 1) / workers;
        std::unordered_map<uint64_t, uint32_t> seen;
        auto &keys = local_keys[w];
        auto &fl = local_flat[w];
        fl.resize(b - a);
        uint64_t last = ~0ull;
        uint32_t last_id = 0;
        for (uint64_t f = a; f < b; f++) {
            const uint64_t k = key_of(f);
            if (k != last) {
                auto it = seen.find(k);
                if (it == seen.end()) {
                    last_id = uint32_t(keys.size());
                    seen.emplace(k, last_id);
                    keys.push_back(k);
                } else {
                    last_id = it->second;
                }
                last = k;
            }
            fl[f - a] = last_id;
        }
    });
    std::unordered_map<uint64_t, uint32_t> global;
    global.emplace(0, 0);
    s.coef_sets.assign({0.f, 0.f});
    std::vector<std::vector<uint32_t>> remap(workers);
    for (unsigned w = 0; w < workers; w++)
        for (uint64_t k : local_keys[w]) {
            auto it = global.find(k);
            uint32_t g;
            if (it == global.end()) {
                g = uint32_t(s.coef_sets.size() / 2);
                global.emplace(k, g);
                float a, b;
                const uint32_t ab = uint32_t(k >> 32), bb = uint32_t(k);
                std::memcpy(&a, &ab, 4);
                std::memcpy(&b, &bb, 4);
                s.coef_sets.push_back(a);
                s.coef_sets.push_back(b);
            } else {
                g = it->second;
            }
            remap[w].push_back(g);
        }
    const bool wide = s.coef_sets.size() / 2 > (narrow ? narrow : 65536);
    s.coef_index.assign(wide ? total : (total + 1) / 2, 0u);
    uint16_t *narrow16 = reinterpret_cast<uint16_t *>(s.coef_index.data());
    parallel_for(workers, [&](size_t w, unsigned) {
        const uint64_t a = total * w / workers;
        const auto &fl = local_flat[w];
        const auto &rm = remap[w];
        if (wide)
            for (size_t q = 0; q < fl.size(); q++) s.coef_index[a + q] = rm[fl[q]];
        else
            for (size_t q = 0; q < fl.size(); q++) narrow16[a + q] = uint16_t(rm[fl[q]]);
    });
}

static void design_map_impl(const fg_sim &s, const fg_sim_desc &d, const uint32_t lo[3], const uint32_t hi[3],
                            std::vector<DesignEdge> &out) {
    const uint64_t *n = s.n;
    const uint64_t nn = s.nn, cells[3] = {n[0] - 1, n[1] - 1, n[2] - 1};
    for (int a = 0; a < 3; a++)
        if (lo[a] >= hi[a] || hi[a] > cells[a]) throw BuildError{"the design box lies outside the grid"};
    const uint64_t box[3] = {hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]};
    std::unordered_set<uint64_t> excluded, debye, seen;
    auto flat = [&](const fg_edge &e) { return uint64_t(e.comp) * nn + node(n, e.at[0], e.at[1], e.at[2]); };
    for (uint32_t l = 0; l < d.n_lumped; l++)
        for (uint32_t q = 0; q < d.lumped[l].n_edges; q++) excluded.insert(flat(d.lumped[l].edges[q]));
    for (uint32_t p = 0; p < d.n_ports; p++) {
        uint64_t o = 0;
        for (uint32_t c = 0; c < d.ports[p].n_columns; c++)
            for (uint32_t q = 0; q < d.ports[p].column_len[c]; q++, o++) excluded.insert(flat(d.ports[p].edges[o]));
    }
    for (uint32_t r = 0; r < d.n_resistive; r++) excluded.insert(uint64_t(d.resistive[r].comp) * nn + d.resistive[r].id);
    for (const fg_debye_edge &e : s.debye) debye.insert(e.f);
    // a thin wire rescales the edges and H nodes around it (sim.rs:556-635), so eps_r enters
    // those edges through another formula: the box keeps clear of every wire by two nodes
    for (uint32_t q = 0; q < d.n_wires; q++) {
        const fg_edge &e = d.wires[q].edge;
        bool near = true;
        for (int a = 0; a < 3; a++) near = near && int64_t(e.at[a]) + 2 >= int64_t(lo[a]) && int64_t(e.at[a]) <= int64_t(hi[a]) + 1;
        if (near) throw BuildError{"the design box touches a thin wire"};
    }
    out.clear();
    static const int offs[4][2] = {{-1, -1}, {0, -1}, {-1, 0}, {0, 0}};
    for (uint64_t ci = lo[0]; ci < hi[0]; ci++)
        for (uint64_t cj = lo[1]; cj < hi[1]; cj++)
            for (uint64_t ck = lo[2]; ck < hi[2]; ck++)
                for (uint32_t c = 0; c < 3; c++) {
                    const uint32_t u = (c + 1) % 3, v = (c + 2) % 3;
                    for (uint32_t du = 0; du < 2; du++)
                        for (uint32_t dv = 0; dv < 2; dv++) {
                            uint64_t p[3] = {ci, cj, ck};
                            p[u] += du;
                            p[v] += dv;
                            const uint64_t f = uint64_t(c) * nn + node(n, p[0], p[1], p[2]);
                            if (!seen.insert(f).second) continue;
                            if (s.cb[f] == 0.f || debye.count(f)) continue;  // eps does not enter
                            if (excluded.count(f)) throw BuildError{"the design box touches a port, lumped or resistive edge"};
                            if (s.ax[u].slot_e[p[u]] >= 0 || s.ax[v].slot_e[p[v]] >= 0)
                                throw BuildError{"the design box reaches into the CPML"};
                            DesignEdge e{f, 0.0, 0.0, {}, {}};
                            double sum_s = 0.0, w_sum = 0.0;
                            for (int q = 0; q < 4; q++) {
                                int64_t cell[3] = {int64_t(p[0]), int64_t(p[1]), int64_t(p[2])};
                                cell[u] += offs[q][0];
                                cell[v] += offs[q][1];
                                const double w = s.ax[u].d[size_t(cell[u])] * s.ax[v].d[size_t(cell[v])];
                                const uint64_t cid = (uint64_t(cell[0]) * cells[1] + uint64_t(cell[1])) * cells[2] + uint64_t(cell[2]);
                                bool in = true;
                                for (int a = 0; a < 3; a++) in = in && uint64_t(cell[a]) >= lo[a] && uint64_t(cell[a]) < hi[a];
                                e.cell[q] = in ? uint32_t(((uint64_t(cell[0]) - lo[0]) * box[1] + (uint64_t(cell[1]) - lo[1])) * box[2] +
                                                          (uint64_t(cell[2]) - lo[2]))
                                               : UINT32_MAX;
                                e.w[q] = w;
                                if (!in) e.fixed += (d.eps ? double(d.eps[cid]) : 1.0) * w;
                                sum_s += (d.sigma ? double(d.sigma[cid]) : 0.0) * w;
                                w_sum += w;
                            }
                            for (int q = 0; q < 4; q++) e.w[q] /= w_sum;
                            e.fixed /= w_sum;
                            e.sig = sum_s / w_sum;
                            out.push_back(e);
                        }
                }
}

std::string design_map(const fg_sim &s, const fg_sim_desc &d, const uint32_t lo[3], const uint32_t hi[3],
                       std::vector<DesignEdge> &out) {
    try {
        design_map_impl(s, d, lo, hi, out);
        return {};
    } catch (const BuildError &e) {
        return e.msg;
    }
}

}  // namespace fg

// ---- C-ABI -------------------------------------------------------------------------------

extern "C" {

int fg_sim_new(const fg_sim_desc *desc, fg_sim **out) {
    if (!desc || !out) return fg::set_error(FG_INVALID, "fg_sim_new: NULL argument");
    fg_sim *s = new fg_sim();
    try {
        fg::build(*s, *desc);
    } catch (const fg::BuildError &e) {
        delete s;
        return fg::set_error(FG_INVALID, e.msg);
    } catch (const std::bad_alloc &) {
        delete s;
        return fg::set_error(FG_TOO_LARGE, "out of host memory building the problem");
    }
    *out = s;
    return FG_OK;
}

void fg_sim_free(fg_sim *sim) { delete sim; }

double fg_sim_dt(const fg_sim *sim) { return sim->dt; }

void fg_sim_dims(const fg_sim *sim, uint32_t n[3]) {
    for (int a = 0; a < 3; a++) n[a] = uint32_t(sim->n[a]);
}

const float *fg_sim_ca(const fg_sim *sim) { return sim->ca.data(); }
const float *fg_sim_cb(const fg_sim *sim) { return sim->cb.data(); }

void fg_sim_counts(const fg_sim *sim, uint64_t counts[4]) {
    counts[0] = sim->debye.size();
    counts[1] = sim->inductors.size();
    counts[2] = sim->sheets.size();
    counts[3] = sim->wires.size();
}

int fg_sim_pack(fg_sim *s, uint64_t narrow, fg_problem *out) {
    if (!s || !out) return fg::set_error(FG_INVALID, "fg_sim_pack: NULL argument");
    const uint64_t *n = s->n;
    const uint64_t nn = s->nn;
    // axes table (run.rs:226-242)
    s->axes.clear();
    for (int a = 0; a < 3; a++) {
        const fg::AxisT &A = s->ax[a];
        for (size_t i = 0; i < A.n; i++) {
            const float row[10] = {A.inv_d[i], A.inv_dd[i], A.raw_inv_d[i], A.raw_inv_dd[i], A.be[i],
                                   A.ce[i],    A.bh[i],     A.ch[i],       float(A.slot_e[i]), float(A.slot_h[i])};
            s->axes.insert(s->axes.end(), row, row + 10);
        }
    }
    fg::coef_sets(*s, narrow);
    // ports, V/I probes, loops, sources (run.rs:302-340)
    s->pk_ports.clear();
    s->pk_edges.clear();
    s->pk_loops.clear();
    s->pk_sources.clear();
    for (size_t q = 0; q < s->ports.size(); q++) {
        const fg_sim::Port &p = s->ports[q];
        const uint32_t axs = p.columns[0][0].comp;
        const auto &mid = p.columns[p.columns.size() / 2];
        fg_port fp{};
        fp.comp = axs;
        fp.first_edge = uint32_t(s->pk_edges.size());
        const double polarity = p.above ? -1.0 : 1.0;
        const double weight = polarity / double(p.columns.size());
        for (auto &col : p.columns)
            for (auto &e : col)
                s->pk_edges.push_back(fg_port_edge{fg::node(n, e.at[0], e.at[1], e.at[2]),
                                                   float(s->ax[axs].d[e.at[axs]] * weight), 0});
        fp.n_edges = uint32_t(s->pk_edges.size()) - fp.first_edge;
        fp.first_loop = uint32_t(s->pk_loops.size());
        double height = 0.0;
        for (auto &e : mid) height += s->ax[axs].d[e.at[axs]];
        for (auto &e : mid) {
            const uint32_t k = e.at[axs];
            const double share = polarity * s->ax[axs].d[k] / height;
            auto terms = fg::port_loop(*s, p, axs, k, share);
            s->pk_loops.insert(s->pk_loops.end(), terms.begin(), terms.end());
        }
        fp.n_loops = uint32_t(s->pk_loops.size()) - fp.first_loop;
        fp.first_source = uint32_t(s->pk_sources.size());
        fp.n_sources = uint32_t(s->port_src[q].size());
        s->pk_sources.insert(s->pk_sources.end(), s->port_src[q].begin(), s->port_src[q].end());
        s->pk_ports.push_back(fp);
    }
    // Debye tables (run.rs:296-301), one block of np pairs per pole set
    size_t np = 0;
    for (auto &ps : s->pole_sets) np = std::max(np, ps.size());
    s->pk_debye_table.clear();
    for (auto &ps : s->pole_sets)
        for (size_t q = 0; q < np; q++) {
            if (q < ps.size()) {
                const double x = ps[q].first, a = ps[q].second;
                const double d = 1.0 + 0.5 * x * s->dt;
                s->pk_debye_table.push_back(float((1.0 - 0.5 * x * s->dt) / d));
                s->pk_debye_table.push_back(float(fg::EPS0 * a * x / d));
            } else {
                s->pk_debye_table.push_back(0.f);
                s->pk_debye_table.push_back(0.f);
            }
        }
    // sheets (run.rs:259-290)
    s->pk_sheets.clear();
    for (auto &sh : s->sheets) {
        if (sh.comp > 1 || sh.id % n[2] == 0) continue;
        const uint64_t hc = sh.comp == 0 ? 1 : 0;
        auto bits = [](uint64_t v) {
            const uint32_t u = uint32_t(v);
            float f;
            std::memcpy(&f, &u, 4);
            return f;
        };
        const float vals[16] = {bits(sh.comp * nn + sh.id), bits(hc * nn + sh.id), bits(hc * nn + sh.id - 1),
                                float(sh.len), float(sh.c), float(sh.r), sh.adaptive ? 1.f : 0.f, 0.f, 0.f, 0.f,
                                0.5f, 0.5f, 0.f, 0.f, bits(sh.faces[0]), bits(sh.faces[1])};
        s->pk_sheets.insert(s->pk_sheets.end(), vals, vals + 16);
    }
    fg_problem p{};
    for (int a = 0; a < 3; a++) p.n[a] = uint32_t(n[a]);
    p.pml = s->pml;
    p.dt = s->dt;
    p.axes = s->axes.data();
    p.coef_wide = s->coef_sets.size() / 2 > (narrow ? narrow : 65536);
    p.n_sets = uint32_t(s->coef_sets.size() / 2);
    p.coef_sets = s->coef_sets.data();
    p.coef_index = s->coef_index.data();
    p.n_ports = uint32_t(s->pk_ports.size());
    p.n_sources = uint32_t(s->pk_sources.size());
    p.n_port_edges = uint32_t(s->pk_edges.size());
    p.n_loops = uint32_t(s->pk_loops.size());
    p.ports = s->pk_ports.data();
    p.port_edges = s->pk_edges.data();
    p.loops = s->pk_loops.data();
    p.sources = s->pk_sources.data();
    p.n_inductors = uint32_t(s->inductors.size());
    p.inductors = s->inductors.data();
    p.n_debye = s->debye.size();
    p.debye_edges = s->debye.data();
    p.n_poles = uint32_t(np);
    p.n_pole_sets = uint32_t(s->pole_sets.size());
    p.debye_table = s->pk_debye_table.data();
    p.n_sheets = s->pk_sheets.size() / 16;
    p.sheets = s->pk_sheets.data();
    p.sheet_states = s->sheet_states;
    p.n_branches = uint32_t(s->sheet_table.size());
    p.branches = s->sheet_table.data();
    p.n_wires = s->wires.size();
    p.wires = s->wires.data();
    *out = p;
    return FG_OK;
}

uint64_t fg_patches(const fg_sim *s, uint32_t inset, uint32_t *gpu, double *pos, double *area) {
    const uint64_t *n = s->n;
    const uint64_t pp = uint64_t(s->pml) + inset;
    if (2 * pp + 2 > std::min({n[0], n[1], n[2]})) return 0;
    const uint64_t lo = pp, hi[3] = {n[0] - 1 - pp, n[1] - 1 - pp, n[2] - 1 - pp};
    uint64_t count = 0;
    for (uint32_t axs = 0; axs < 3; axs++) {
        const uint32_t u = (axs + 1) % 3, v = (axs + 2) % 3;
        for (int side = 0; side < 2; side++) {
            const uint64_t plane = side ? hi[axs] : lo;
            for (uint64_t a = lo; a < hi[u]; a++)
                for (uint64_t b = lo; b < hi[v]; b++, count++) {
                    if (pos) {
                        double *ps = pos + 3 * count;
                        ps[axs] = s->line[axs][plane];
                        ps[u] = 0.5 * (s->line[u][a] + s->line[u][a + 1]);
                        ps[v] = 0.5 * (s->line[v][b] + s->line[v][b + 1]);
                    }
                    if (area) area[count] = s->ax[u].d[a] * s->ax[v].d[b];
                    if (!gpu) continue;
                    uint64_t q[3];
                    q[axs] = plane, q[u] = a, q[v] = b;
                    auto at = [&](int da, int du, int dv) {
                        uint64_t r[3] = {q[0], q[1], q[2]};
                        r[axs] += da, r[u] += du, r[v] += dv;
                        return uint32_t(fg::node(n, r[0], r[1], r[2]));
                    };
                    uint32_t *g = gpu + 13 * count;
                    g[0] = axs | (side ? 4u : 0u);
                    g[1] = at(0, 0, 0), g[2] = at(0, 0, 1), g[3] = at(0, 0, 0), g[4] = at(0, 1, 0);
                    g[5] = at(0, 0, 0), g[6] = at(-1, 0, 0), g[7] = at(0, 1, 0), g[8] = at(-1, 1, 0);
                    g[9] = at(0, 0, 0), g[10] = at(-1, 0, 0), g[11] = at(0, 0, 1), g[12] = at(-1, 0, 1);
                }
        }
    }
    return count;
}

}  // extern "C"
