// lammps_subset.cpp -- deck reader and host force terms (see lammps_subset.hpp).
#include "lammps_subset.hpp"

#include <algorithm>
#include <array>
#include <cmath>
#include <fstream>
#include <set>
#include <sstream>
#include <stdexcept>

namespace selm_host {

namespace {

std::vector<std::string> split(const std::string &s) {
    std::istringstream is(s);
    std::vector<std::string> out;
    std::string t;
    while (is >> t) out.push_back(t);
    return out;
}

std::string strip_comment(const std::string &s) {
    bool q = false;
    for (size_t i = 0; i < s.size(); i++) {
        if (s[i] == '"' || s[i] == '\'') q = !q;
        if (s[i] == '#' && !q) return s.substr(0, i);
    }
    return s;
}

double to_d(const std::string &s, const std::string &ctx) {
    try {
        size_t n = 0;
        double v = std::stod(s, &n);
        if (n != s.size()) throw std::invalid_argument(s);
        return v;
    } catch (const std::exception &) {
        throw std::runtime_error("expected a number, got '" + s + "' in: " + ctx);
    }
}

long to_i(const std::string &s, const std::string &ctx) {
    try {
        size_t n = 0;
        long v = std::stol(s, &n);
        if (n != s.size()) throw std::invalid_argument(s);
        return v;
    } catch (const std::exception &) {
        throw std::runtime_error("expected an integer, got '" + s + "' in: " + ctx);
    }
}

// Type range of a coefficient command: "3", "*", "2*", "*3", "2*4".
void type_range(const std::string &s, int n, int &lo, int &hi, const std::string &ctx) {
    const size_t st = s.find('*');
    if (st == std::string::npos) {
        lo = hi = (int)to_i(s, ctx);
    } else {
        lo = st == 0 ? 1 : (int)to_i(s.substr(0, st), ctx);
        hi = st + 1 == s.size() ? n : (int)to_i(s.substr(st + 1), ctx);
    }
    if (lo < 1 || hi > n || lo > hi) throw std::runtime_error("type out of range in: " + ctx);
}

struct Script {
    std::map<std::string, std::string> vars;
    std::string dir;

    std::string substitute(const std::string &line) const {
        std::string out;
        for (size_t i = 0; i < line.size(); i++) {
            if (line[i] != '$' || i + 1 >= line.size()) {
                out += line[i];
                continue;
            }
            std::string name;
            if (line[i + 1] == '{') {
                const size_t e = line.find('}', i + 2);
                if (e == std::string::npos) throw std::runtime_error("unterminated ${ in: " + line);
                name = line.substr(i + 2, e - i - 2);
                i = e;
            } else {
                name = std::string(1, line[i + 1]);
                i += 1;
            }
            auto it = vars.find(name);
            if (it == vars.end()) throw std::runtime_error("undefined variable '" + name + "' in: " + line);
            out += it->second;
        }
        return out;
    }
};

void read_data(Deck &d, const std::string &path, const std::string &atom_style) {
    std::ifstream in(path);
    if (!in) throw std::runtime_error("cannot open data file " + path);
    std::string line;
    std::getline(in, line);  // the first line is a title
    int nbonds = 0, nangles = 0;
    // header: counts and box until the first section keyword
    std::string section;
    while (std::getline(in, line)) {
        const std::string s = strip_comment(line);
        auto t = split(s);
        if (t.empty()) continue;
        if (t.size() >= 2 && t[1] == "atoms") d.natoms = (int)to_i(t[0], s);
        else if (t.size() >= 2 && t[1] == "bonds") nbonds = (int)to_i(t[0], s);
        else if (t.size() >= 2 && t[1] == "angles") nangles = (int)to_i(t[0], s);
        else if (t.size() >= 2 && (t[1] == "dihedrals" || t[1] == "impropers")) {
            if (to_i(t[0], s) != 0) throw std::runtime_error("dihedrals/impropers are not supported: " + s);
        } else if (t.size() >= 3 && t[1] == "atom" && t[2] == "types") d.ntypes = (int)to_i(t[0], s);
        else if (t.size() >= 3 && t[1] == "bond" && t[2] == "types") d.nbond_types = (int)to_i(t[0], s);
        else if (t.size() >= 3 && t[1] == "angle" && t[2] == "types") d.nangle_types = (int)to_i(t[0], s);
        else if (t.size() >= 3 && (t[1] == "dihedral" || t[1] == "improper") && t[2] == "types") {
        } else if (t.size() == 4 && t[2] == "xlo") { d.lo[0] = to_d(t[0], s); d.hi[0] = to_d(t[1], s); }
        else if (t.size() == 4 && t[2] == "ylo") { d.lo[1] = to_d(t[0], s); d.hi[1] = to_d(t[1], s); }
        else if (t.size() == 4 && t[2] == "zlo") { d.lo[2] = to_d(t[0], s); d.hi[2] = to_d(t[1], s); }
        else if (t.size() == 6 && t[3] == "xy") throw std::runtime_error("triclinic boxes are not supported");
        else { section = s; break; }
    }
    const int n = d.natoms;
    d.x.assign(3 * (size_t)n, 0.0);
    d.v.assign(3 * (size_t)n, 0.0);
    d.id.assign(n, 0);
    d.mol.assign(n, 0);
    d.type.assign(n, 0);
    d.mass_of_type.assign(d.ntypes + 1, 0.0);
    std::vector<int> raw_id(n), raw_mol(n), raw_type(n);
    std::vector<double> raw_x(3 * (size_t)n);
    std::vector<std::array<double, 3>> vel_val;
    std::vector<long> vel_id;
    std::vector<long> b_type, b_i, b_j, a_type, a_i, a_j, a_k;
    bool have_atoms = false;

    auto next_data_line = [&](std::vector<std::string> &t, const std::string &what) {
        while (std::getline(in, line)) {
            t = split(strip_comment(line));
            if (!t.empty()) return;
        }
        throw std::runtime_error("data file ended inside " + what);
    };

    while (!section.empty()) {
        const auto st = split(section);
        const std::string key = st[0];
        std::vector<std::string> t;
        if (key == "Masses") {
            for (int i = 0; i < d.ntypes; i++) {
                next_data_line(t, "Masses");
                const int ty = (int)to_i(t.at(0), line);
                if (ty < 1 || ty > d.ntypes) throw std::runtime_error("bad type in Masses: " + line);
                d.mass_of_type[ty] = to_d(t.at(1), line);
            }
        } else if (key == "Atoms") {
            for (int i = 0; i < n; i++) {
                next_data_line(t, "Atoms");
                int c = 0;
                raw_id[i] = (int)to_i(t.at(c++), line);
                if (atom_style == "atomic") raw_mol[i] = 0;
                else raw_mol[i] = (int)to_i(t.at(c++), line);
                raw_type[i] = (int)to_i(t.at(c++), line);
                if (atom_style == "full" || atom_style == "charge") c++;  // charge column
                for (int a = 0; a < 3; a++) raw_x[3 * i + a] = to_d(t.at(c++), line);
            }
            have_atoms = true;
        } else if (key == "Velocities") {
            for (int i = 0; i < n; i++) {
                next_data_line(t, "Velocities");
                vel_id.push_back(to_i(t.at(0), line));
                vel_val.push_back({to_d(t.at(1), line), to_d(t.at(2), line), to_d(t.at(3), line)});
            }
        } else if (key == "Bonds") {
            for (int i = 0; i < nbonds; i++) {
                next_data_line(t, "Bonds");
                b_type.push_back(to_i(t.at(1), line));
                b_i.push_back(to_i(t.at(2), line));
                b_j.push_back(to_i(t.at(3), line));
            }
        } else if (key == "Angles") {
            for (int i = 0; i < nangles; i++) {
                next_data_line(t, "Angles");
                a_type.push_back(to_i(t.at(1), line));
                a_i.push_back(to_i(t.at(2), line));
                a_j.push_back(to_i(t.at(3), line));
                a_k.push_back(to_i(t.at(4), line));
            }
        } else {
            throw std::runtime_error("unsupported data file section: " + section);
        }
        section.clear();
        while (std::getline(in, line)) {
            const std::string s = strip_comment(line);
            if (!split(s).empty()) {
                section = s;
                break;
            }
        }
    }
    if (!have_atoms) throw std::runtime_error("data file has no Atoms section");
    // ascending atom-ID order
    std::vector<int> ord(n);
    for (int i = 0; i < n; i++) ord[i] = i;
    std::sort(ord.begin(), ord.end(), [&](int a, int b) { return raw_id[a] < raw_id[b]; });
    std::map<long, int> index;
    for (int k = 0; k < n; k++) {
        const int i = ord[k];
        d.id[k] = raw_id[i];
        d.mol[k] = raw_mol[i];
        d.type[k] = raw_type[i];
        if (d.type[k] < 1 || d.type[k] > d.ntypes) throw std::runtime_error("atom type out of range");
        for (int a = 0; a < 3; a++) d.x[3 * k + a] = raw_x[3 * i + a];
        if (!index.emplace(raw_id[i], k).second) throw std::runtime_error("duplicate atom ID");
    }
    auto at = [&](long aid) {
        auto it = index.find(aid);
        if (it == index.end()) throw std::runtime_error("unknown atom ID " + std::to_string(aid));
        return it->second;
    };
    for (size_t q = 0; q < vel_id.size(); q++)
        for (int a = 0; a < 3; a++) d.v[3 * at(vel_id[q]) + a] = vel_val[q][a];
    for (size_t q = 0; q < b_type.size(); q++) {
        if (b_type[q] < 1 || b_type[q] > d.nbond_types) throw std::runtime_error("bond type out of range");
        d.bond_type.push_back((int)b_type[q]);
        d.bond_i.push_back(at(b_i[q]));
        d.bond_j.push_back(at(b_j[q]));
    }
    for (size_t q = 0; q < a_type.size(); q++) {
        if (a_type[q] < 1 || a_type[q] > d.nangle_types) throw std::runtime_error("angle type out of range");
        d.angle_type.push_back((int)a_type[q]);
        d.angle_i.push_back(at(a_i[q]));
        d.angle_j.push_back(at(a_j[q]));
        d.angle_k.push_back(at(a_k[q]));
    }
    if ((int)d.bond_type.size() != nbonds || (int)d.angle_type.size() != nangles)
        throw std::runtime_error("data file: Bonds/Angles sections do not match the header counts");
    d.fext.assign(3 * (size_t)n, 0.0);
    d.bond_K.assign(d.nbond_types + 1, NAN);
    d.bond_r0.assign(d.nbond_types + 1, NAN);
    d.angle_K.assign(d.nangle_types + 1, NAN);
    d.angle_theta0.assign(d.nangle_types + 1, NAN);
    const size_t T = (size_t)(d.ntypes + 1);
    d.lj_eps.assign(T * T, 0.0);
    d.lj_sigma.assign(T * T, 0.0);
    d.lj_rc.assign(T * T, 0.0);
    d.lj_set.assign(T * T, 0);
    std::vector<int> all(n);
    for (int i = 0; i < n; i++) all[i] = i;
    d.groups["all"] = all;
}

std::vector<int> parse_ids(const std::vector<std::string> &t, size_t from, const std::string &ctx) {
    std::vector<int> out;
    for (size_t k = from; k < t.size(); k++) {
        const size_t c = t[k].find(':');
        if (c == std::string::npos) {
            out.push_back((int)to_i(t[k], ctx));
        } else {
            const long a = to_i(t[k].substr(0, c), ctx), b = to_i(t[k].substr(c + 1), ctx);
            for (long q = a; q <= b; q++) out.push_back((int)q);
        }
    }
    return out;
}

}  // namespace

std::vector<double> Deck::masses() const {
    std::vector<double> m(natoms);
    for (int i = 0; i < natoms; i++) m[i] = mass_of_type[type[i]];
    return m;
}

bool Deck::has_host_terms() const { return !angle_type.empty() || !pair_style.empty(); }

Deck load_deck(const std::string &script_path) {
    std::ifstream in(script_path);
    if (!in) throw std::runtime_error("cannot open " + script_path);
    Script sc;
    const size_t slash = script_path.find_last_of('/');
    sc.dir = slash == std::string::npos ? "" : script_path.substr(0, slash + 1);
    Deck d;
    std::string atom_style = "atomic";
    bool have_data = false;
    std::string mix = "geometric";
    std::string raw, line;
    while (std::getline(in, raw)) {
        line = strip_comment(raw);
        while (!line.empty() && (line.back() == ' ' || line.back() == '\t' || line.back() == '\r')) line.pop_back();
        if (!line.empty() && line.back() == '&') {  // continuation
            line.pop_back();
            std::string more;
            if (std::getline(in, more)) {
                raw = line + " " + more;
                line = strip_comment(raw);
            }
        }
        auto t0 = split(line);
        if (t0.empty()) continue;
        const bool is_var = t0[0] == "variable";
        const std::string cmdline = is_var ? line : sc.substitute(line);
        auto t = split(cmdline);
        const std::string &c = t[0];
        auto need = [&](size_t k) {
            if (t.size() < k) throw std::runtime_error("too few arguments: " + cmdline);
        };
        if (c == "variable") {
            need(4);
            const std::string &style = t[2];
            if (style != "equal" && style != "index" && style != "string" && style != "universe")
                throw std::runtime_error("unsupported variable style: " + cmdline);
            // equal-style variables are replaced; index/string/universe keep their first definition
            if (style == "equal" || !sc.vars.count(t[1])) sc.vars[t[1]] = sc.substitute(t[3]);
        } else if (c == "units") {
            need(2);
            d.units = t[1];
        } else if (c == "boundary") {
            need(4);
            if (t[1] != "p" || t[2] != "p" || t[3] != "p")
                throw std::runtime_error("only fully periodic boundaries are supported: " + cmdline);
        } else if (c == "atom_style") {
            need(2);
            atom_style = t[1];
            if (atom_style != "angle" && atom_style != "bond" && atom_style != "molecular" && atom_style != "full" &&
                atom_style != "atomic")
                throw std::runtime_error("unsupported atom_style: " + cmdline);
        } else if (c == "log" || c == "thermo" || c == "thermo_style" || c == "thermo_modify" || c == "neighbor" ||
                   c == "neigh_modify" || c == "atom_modify" || c == "comm_modify" || c == "newton" ||
                   c == "processors" || c == "echo" || c == "dump_modify" || c == "print") {
            // no effect on the SELM step
        } else if (c == "read_data") {
            need(2);
            const std::string p = (t[1].size() && t[1][0] == '/') ? t[1] : sc.dir + t[1];
            read_data(d, p, atom_style);
            have_data = true;
        } else if (c == "mass") {
            need(3);
            if (!have_data) throw std::runtime_error("mass before read_data");
            int lo, hi;
            type_range(t[1], d.ntypes, lo, hi, cmdline);
            for (int q = lo; q <= hi; q++) d.mass_of_type[q] = to_d(t[2], cmdline);
        } else if (c == "bond_style") {
            need(2);
            if (t[1] != "harmonic") throw std::runtime_error("unsupported bond_style: " + cmdline);
            d.bond_style = t[1];
        } else if (c == "angle_style") {
            need(2);
            if (t[1] != "harmonic") throw std::runtime_error("unsupported angle_style: " + cmdline);
            d.angle_style = t[1];
        } else if (c == "pair_style") {
            need(2);
            if (t[1] == "none") {
                d.pair_style.clear();
            } else if (t[1] == "lj/cut") {
                need(3);
                d.pair_style = t[1];
                d.pair_rc = to_d(t[2], cmdline);
            } else {
                throw std::runtime_error("unsupported pair_style: " + cmdline);
            }
        } else if (c == "pair_modify") {
            for (size_t k = 1; k + 1 < t.size(); k++)
                if (t[k] == "mix") mix = t[k + 1];
            if (mix != "geometric" && mix != "arithmetic") throw std::runtime_error("unsupported mix: " + cmdline);
        } else if (c == "special_bonds") {
            for (size_t k = 1; k < t.size(); k++) {
                if (t[k] == "lj" || t[k] == "lj/coul") {
                    if (k + 3 >= t.size()) throw std::runtime_error("special_bonds lj needs three weights: " + cmdline);
                    for (int q = 0; q < 3; q++) d.special_lj[q] = to_d(t.at(k + 1 + q), cmdline);
                    k += 3;
                } else if (t[k] == "coul") {
                    if (k + 3 >= t.size()) throw std::runtime_error("special_bonds coul needs three weights: " + cmdline);
                    k += 3;
                } else {
                    throw std::runtime_error("unsupported special_bonds keyword: " + cmdline);
                }
            }
        } else if (c == "bond_coeff") {
            need(4);
            if (!have_data) throw std::runtime_error("bond_coeff before read_data");
            int lo, hi;
            type_range(t[1], d.nbond_types, lo, hi, cmdline);
            for (int q = lo; q <= hi; q++) {
                d.bond_K[q] = to_d(t[2], cmdline);
                d.bond_r0[q] = to_d(t[3], cmdline);
            }
        } else if (c == "angle_coeff") {
            need(4);
            if (!have_data) throw std::runtime_error("angle_coeff before read_data");
            int lo, hi;
            type_range(t[1], d.nangle_types, lo, hi, cmdline);
            for (int q = lo; q <= hi; q++) {
                d.angle_K[q] = to_d(t[2], cmdline);
                d.angle_theta0[q] = to_d(t[3], cmdline) * M_PI / 180.0;
            }
        } else if (c == "pair_coeff") {
            need(5);
            if (!have_data) throw std::runtime_error("pair_coeff before read_data");
            if (d.pair_style != "lj/cut") throw std::runtime_error("pair_coeff without pair_style lj/cut");
            int ilo, ihi, jlo, jhi;
            type_range(t[1], d.ntypes, ilo, ihi, cmdline);
            type_range(t[2], d.ntypes, jlo, jhi, cmdline);
            const double eps = to_d(t[3], cmdline), sig = to_d(t[4], cmdline);
            const double rc = t.size() > 5 ? to_d(t[5], cmdline) : d.pair_rc;
            const size_t T = (size_t)d.ntypes + 1;
            for (int i = ilo; i <= ihi; i++)
                for (int j = std::max(jlo, i); j <= jhi; j++) {
                    for (int s = 0; s < 2; s++) {
                        const size_t k = s ? (size_t)j * T + i : (size_t)i * T + j;
                        d.lj_eps[k] = eps;
                        d.lj_sigma[k] = sig;
                        d.lj_rc[k] = rc;
                        d.lj_set[k] = 1;
                    }
                }
        } else if (c == "group") {
            need(4);
            if (!have_data) throw std::runtime_error("group before read_data");
            const std::vector<int> ids = parse_ids(t, 3, cmdline);
            std::set<int> want(ids.begin(), ids.end());
            std::vector<int> members;
            for (int i = 0; i < d.natoms; i++) {
                const int key = t[2] == "id" ? d.id[i] : t[2] == "type" ? d.type[i] : t[2] == "molecule" ? d.mol[i] : -1;
                if (t[2] != "id" && t[2] != "type" && t[2] != "molecule")
                    throw std::runtime_error("unsupported group style: " + cmdline);
                if (want.count(key)) members.push_back(i);
            }
            d.groups[t[1]] = members;
        } else if (c == "fix") {
            need(4);
            if (t[3] != "addforce") throw std::runtime_error("unsupported fix (only addforce): " + cmdline);
            need(7);
            auto g = d.groups.find(t[2]);
            if (g == d.groups.end()) throw std::runtime_error("unknown group: " + cmdline);
            for (int i : g->second)
                for (int a = 0; a < 3; a++) d.fext[3 * (size_t)i + a] += to_d(t[4 + a], cmdline);
        } else if (c == "timestep") {
            need(2);
            d.dt = to_d(t[1], cmdline);
            d.dt_set = true;
        } else if (c == "dump") {
            need(6);
            if (t[3] != "dcd") throw std::runtime_error("unsupported dump style (only dcd): " + cmdline);
            d.dumps.push_back({(int)to_i(t[4], cmdline), t[5]});
        } else {
            throw std::runtime_error("unsupported command: " + cmdline);
        }
    }
    if (!have_data) throw std::runtime_error("deck has no read_data");
    if (!d.dt_set) {
        // LAMMPS's default timestep for the unit style
        d.dt = d.units == "real" ? 1.0 : d.units == "metal" ? 0.001 : 0.005;
    }
    for (size_t q = 0; q < d.bond_type.size(); q++)
        if (std::isnan(d.bond_K[d.bond_type[q]])) throw std::runtime_error("bond type without bond_coeff");
    for (size_t q = 0; q < d.angle_type.size(); q++)
        if (std::isnan(d.angle_K[d.angle_type[q]])) throw std::runtime_error("angle type without angle_coeff");
    if (!d.bond_type.empty() && d.bond_style.empty()) throw std::runtime_error("bonds without bond_style");
    if (!d.angle_type.empty() && d.angle_style.empty()) throw std::runtime_error("angles without angle_style");
    if (d.pair_style == "lj/cut") {
        const size_t T = (size_t)d.ntypes + 1;
        for (int i = 1; i <= d.ntypes; i++)
            for (int j = 1; j <= d.ntypes; j++) {
                const size_t k = (size_t)i * T + j;
                if (d.lj_set[k]) continue;
                const size_t ii = (size_t)i * T + i, jj = (size_t)j * T + j;
                if (!d.lj_set[ii] || !d.lj_set[jj]) throw std::runtime_error("pair coefficients not set for all types");
                d.lj_eps[k] = std::sqrt(d.lj_eps[ii] * d.lj_eps[jj]);
                if (mix == "geometric") {
                    d.lj_sigma[k] = std::sqrt(d.lj_sigma[ii] * d.lj_sigma[jj]);
                    d.lj_rc[k] = std::sqrt(d.lj_rc[ii] * d.lj_rc[jj]);
                } else {
                    d.lj_sigma[k] = 0.5 * (d.lj_sigma[ii] + d.lj_sigma[jj]);
                    d.lj_rc[k] = 0.5 * (d.lj_rc[ii] + d.lj_rc[jj]);
                }
            }
    }
    return d;
}

namespace {

inline void min_image(const Deck &d, double *del) {
    for (int a = 0; a < 3; a++) {
        const double L = d.hi[a] - d.lo[a];
        if (std::fabs(del[a]) > 0.5 * L) del[a] += del[a] < 0.0 ? L : -L;
    }
}

// special-bond weight of pair (i, j): 1-2, 1-3, 1-4 partners through bonds
struct Special {
    std::vector<std::vector<int>> nb;  // bonded partners
    explicit Special(const Deck &d) : nb(d.natoms) {
        for (size_t q = 0; q < d.bond_i.size(); q++) {
            nb[d.bond_i[q]].push_back(d.bond_j[q]);
            nb[d.bond_j[q]].push_back(d.bond_i[q]);
        }
    }
    // 0 = not special, 1 = 1-2, 2 = 1-3, 3 = 1-4
    int order(int i, int j) const {
        for (int a : nb[i])
            if (a == j) return 1;
        for (int a : nb[i])
            for (int b : nb[a])
                if (b == j && b != i) return 2;
        for (int a : nb[i])
            for (int b : nb[a]) {
                if (b == i) continue;
                for (int c : nb[b])
                    if (c == j && c != a && c != i) return 3;
            }
        return 0;
    }
};

}  // namespace

double host_forces(const Deck &d, const double *x, double *f) {
    double energy = 0.0;
    // harmonic angles: E = K (theta - theta0)^2 with theta the angle at atom j.
    // dE/dx_i = 2 K dtheta * dtheta/dcos * dcos/dx_i, dtheta/dcos = -1/sin theta;
    // sin theta is held away from 0 by 0.001 as in LAMMPS's angle_harmonic.
    constexpr double kSinFloor = 0.001;
    for (size_t q = 0; q < d.angle_type.size(); q++) {
        const int i = d.angle_i[q], j = d.angle_j[q], k = d.angle_k[q];
        double a[3], b[3];
        for (int c = 0; c < 3; c++) {
            a[c] = x[3 * i + c] - x[3 * j + c];
            b[c] = x[3 * k + c] - x[3 * j + c];
        }
        min_image(d, a);
        min_image(d, b);
        const double ra2 = a[0] * a[0] + a[1] * a[1] + a[2] * a[2];
        const double rb2 = b[0] * b[0] + b[1] * b[1] + b[2] * b[2];
        const double ra = std::sqrt(ra2), rb = std::sqrt(rb2);
        double cs = (a[0] * b[0] + a[1] * b[1] + a[2] * b[2]) / (ra * rb);
        cs = std::min(1.0, std::max(-1.0, cs));
        double sn = std::sqrt(1.0 - cs * cs);
        if (sn < kSinFloor) sn = kSinFloor;
        const double dth = std::acos(cs) - d.angle_theta0[d.angle_type[q]];
        const double K = d.angle_K[d.angle_type[q]];
        energy += K * dth * dth;
        const double g = 2.0 * K * dth / sn;  // -dE/dcos
        for (int c = 0; c < 3; c++) {
            const double fi = g * (b[c] / (ra * rb) - cs * a[c] / ra2);
            const double fk = g * (a[c] / (ra * rb) - cs * b[c] / rb2);
            f[3 * i + c] += fi;
            f[3 * k + c] += fk;
            f[3 * j + c] -= fi + fk;
        }
    }
    if (d.pair_style == "lj/cut" && d.natoms > 1) {
        const size_t T = (size_t)d.ntypes + 1;
        double rcmax = 0.0;
        for (double r : d.lj_rc) rcmax = std::max(rcmax, r);
        double L[3];
        for (int a = 0; a < 3; a++) L[a] = d.hi[a] - d.lo[a];
        if (2.0 * rcmax >= std::min(L[0], std::min(L[1], L[2])))
            throw std::runtime_error("lj/cut cutoff must be below half the box edge");
        const Special sp(d);
        // cell list with cells no smaller than the largest cutoff
        int nc[3];
        for (int a = 0; a < 3; a++) nc[a] = std::max(1, (int)std::floor(L[a] / rcmax));
        const bool use_cells = nc[0] >= 3 && nc[1] >= 3 && nc[2] >= 3;
        std::vector<std::vector<int>> cells;
        auto cell_of = [&](int i, int *c) {
            for (int a = 0; a < 3; a++) {
                double s = (x[3 * i + a] - d.lo[a]) / L[a];
                s -= std::floor(s);
                c[a] = std::min(nc[a] - 1, (int)(s * nc[a]));
            }
        };
        auto pair = [&](int i, int j) {
            double del[3];
            for (int c = 0; c < 3; c++) del[c] = x[3 * i + c] - x[3 * j + c];
            min_image(d, del);
            const double r2 = del[0] * del[0] + del[1] * del[1] + del[2] * del[2];
            const size_t k = (size_t)d.type[i] * T + d.type[j];
            const double rc = d.lj_rc[k];
            if (r2 >= rc * rc) return;
            const int o = sp.order(i, j);
            const double w = o == 0 ? 1.0 : d.special_lj[o - 1];
            if (w == 0.0) return;
            const double s2 = d.lj_sigma[k] * d.lj_sigma[k] / r2;
            const double s6 = s2 * s2 * s2;
            const double eps = d.lj_eps[k];
            energy += w * 4.0 * eps * (s6 * s6 - s6);
            const double fr = w * 24.0 * eps * (2.0 * s6 * s6 - s6) / r2;  // F = fr * del
            for (int c = 0; c < 3; c++) {
                f[3 * i + c] += fr * del[c];
                f[3 * j + c] -= fr * del[c];
            }
        };
        if (!use_cells) {
            for (int i = 0; i < d.natoms; i++)
                for (int j = i + 1; j < d.natoms; j++) pair(i, j);
        } else {
            cells.assign((size_t)nc[0] * nc[1] * nc[2], {});
            for (int i = 0; i < d.natoms; i++) {
                int c[3];
                cell_of(i, c);
                cells[((size_t)c[2] * nc[1] + c[1]) * nc[0] + c[0]].push_back(i);
            }
            for (int i = 0; i < d.natoms; i++) {
                int c[3];
                cell_of(i, c);
                for (int dz = -1; dz <= 1; dz++)
                    for (int dy = -1; dy <= 1; dy++)
                        for (int dx = -1; dx <= 1; dx++) {
                            const int cx = (c[0] + dx + nc[0]) % nc[0], cy = (c[1] + dy + nc[1]) % nc[1],
                                      cz = (c[2] + dz + nc[2]) % nc[2];
                            for (int j : cells[((size_t)cz * nc[1] + cy) * nc[0] + cx])
                                if (j > i) pair(i, j);
                        }
            }
        }
    }
    return energy;
}

double all_forces(const Deck &d, const double *x, double *f) {
    std::fill(f, f + 3 * (size_t)d.natoms, 0.0);
    double energy = 0.0;
    for (size_t q = 0; q < d.bond_type.size(); q++) {
        const int i = d.bond_i[q], j = d.bond_j[q];
        double del[3];
        for (int c = 0; c < 3; c++) del[c] = x[3 * i + c] - x[3 * j + c];
        min_image(d, del);
        const double r = std::sqrt(del[0] * del[0] + del[1] * del[1] + del[2] * del[2]);
        const double dr = r - d.bond_r0[d.bond_type[q]];
        const double K = d.bond_K[d.bond_type[q]];
        energy += K * dr * dr;
        const double fb = r > 0.0 ? -2.0 * K * dr / r : 0.0;
        for (int c = 0; c < 3; c++) {
            f[3 * i + c] += fb * del[c];
            f[3 * j + c] -= fb * del[c];
        }
    }
    energy += host_forces(d, x, f);
    for (size_t k = 0; k < 3 * (size_t)d.natoms; k++) f[k] += d.fext[k];
    return energy;
}

}  // namespace selm_host
