// Model readers (native JSON, MPS), result JSON and solution files.
//
// The file formats and their rules are Niryukti's (src/io.cpp): the native
// JSON schema (:84-140), free- and fixed-format MPS with RANGES, BOUNDS and
// OBJSENSE (:299-585), and the result document keys (:902-995), so models and
// results move between the two engines unchanged.  Integer variables and
// quadratic objectives are rejected: this engine solves LPs.
#include "io.hpp"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <set>
#include <sstream>
#include <stdexcept>
#include <unordered_map>

namespace nyk {

std::string engine_version() { return "0.2.2+mi355x"; }

namespace {

std::string trim(const std::string &s) {
    auto a = s.find_first_not_of(" \t\r\n");
    return a == std::string::npos ? "" : s.substr(a, s.find_last_not_of(" \t\r\n") - a + 1);
}
std::string upper(std::string s) {
    for (auto &c : s) c = char(std::toupper(static_cast<unsigned char>(c)));
    return s;
}
std::vector<std::string> words(const std::string &s) {
    std::istringstream in(s);
    std::vector<std::string> t;
    std::string a;
    while (in >> a) t.push_back(a);
    return t;
}
// Finite number; Fortran 'D' exponents accepted (src/io.cpp:32-41).
double number(std::string s) {
    for (auto &c : s)
        if (c == 'D' || c == 'd') c = 'e';
    size_t end = 0;
    double v;
    try {
        v = std::stod(s, &end);
    } catch (const std::exception &) {
        throw std::runtime_error("Invalid finite number: " + s);
    }
    if (end != s.size() || !std::isfinite(v)) throw std::runtime_error("Invalid finite number: " + s);
    return v;
}
double sense_of(std::string s) {
    s = upper(s);
    if (s == "MIN" || s == "MINIMIZE") return 1;
    if (s == "MAX" || s == "MAXIMIZE") return -1;
    throw std::runtime_error("Unknown objective sense");
}
// Objective sense folded into the costs: canonical minimisation (src/io.cpp:74-83).
void canonicalize(Model &m) {
    for (double &v : m.c) v *= m.sense;
    m.offset *= m.sense;
    m.validate();
}
std::string slurp(const std::string &path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("Cannot open model: " + path);
    std::ostringstream ss;
    ss << f.rdbuf();
    return ss.str();
}
// Bound field of the native JSON format: null -> fallback, "inf"/"-inf" strings.
double json_bound(const Json *j, double fallback) {
    if (!j || j->is_null()) return fallback;
    if (j->is_string()) {
        auto s = upper(j->str());
        if (s == "INF" || s == "+INF") return kInf;
        if (s == "-INF") return -kInf;
        throw std::runtime_error("Invalid bound string");
    }
    return j->number();
}
}  // namespace

Model read_json_model(const std::string &text) {
    Json j = Json::parse(text);
    Model m;
    if (const Json *n = j.find("name")) m.name = n->str();
    if (const Json *s = j.find("sense")) m.sense = sense_of(s->str());
    std::unordered_map<std::string, int64_t> ids;
    const Json &vars = j.at("variables");
    for (size_t k = 0; k < vars.size(); ++k) {
        const Json &v = vars[k];
        std::string name = v.at("name").str();
        if (ids.count(name)) throw std::runtime_error("Duplicate variable");
        ids[name] = int64_t(m.names.size());
        m.names.push_back(name);
        const Json *lb = v.find("lb");
        static const Json zero(0.0);
        m.lb.push_back(json_bound(lb ? lb : &zero, -kInf));
        m.ub.push_back(json_bound(v.find("ub"), kInf));
        if (const Json *t = v.find("type")) {
            const auto &type = t->str();
            if (type == "integer" || type == "binary")
                throw std::runtime_error("UNSUPPORTED: integer variables are outside this LP engine");
            if (type != "continuous") throw std::runtime_error("Unknown variable type");
        }
    }
    const Json &obj = j.at("objective");
    m.c = obj.at("linear").numbers();
    for (const char *key : {"quadratic_diagonal", "quadratic_sparse", "quadratic"})
        if (const Json *q = obj.find(key))
            if (!(q->is_array() && (q->size() == 0 || (std::string(key) == "quadratic_diagonal" &&
                                                        [&] {
                                                            for (double d : q->numbers())
                                                                if (d != 0) return false;
                                                            return true;
                                                        }()))))
                throw std::runtime_error("UNSUPPORTED: quadratic objectives are outside this LP engine");
    if (const Json *off = obj.find("offset")) m.offset = off->number();
    std::vector<Entry> e;
    const Json &rows = j.at("constraints");
    for (size_t i = 0; i < rows.size(); ++i) {
        const Json &row = rows[i];
        const Json *name = row.find("name");
        m.row_names.push_back(name ? name->str() : "r" + std::to_string(i));
        m.rl.push_back(json_bound(row.find("lb"), -kInf));
        m.ru.push_back(json_bound(row.find("ub"), kInf));
        for (const auto &kv : row.at("coefficients").items()) {
            auto it = ids.find(kv.first);
            if (it == ids.end()) throw std::runtime_error("Unknown variable: " + kv.first);
            e.push_back({int64_t(i), it->second, kv.second.number()});
        }
    }
    m.A = Csr::from_entries(int64_t(m.rl.size()), int64_t(m.names.size()), std::move(e));
    canonicalize(m);
    return m;
}

Model read_mps_model(const std::string &text) {
    Model m;
    std::string section, obj, rhs_set, range_set, bound_set, last_col;
    std::unordered_map<std::string, int64_t> vars, rows;
    std::vector<char> rowtype;
    std::vector<double> rhs, ranges;
    std::vector<bool> has_range;
    std::vector<Entry> e;
    std::set<std::pair<std::string, std::string>> bound_seen;
    bool ended = false;
    int lineno = 0;
    auto var = [&](const std::string &s) {
        auto it = vars.find(s);
        if (it != vars.end()) return it->second;
        int64_t k = int64_t(m.names.size());
        vars.emplace(s, k);
        m.names.push_back(s);
        m.c.push_back(0);
        m.lb.push_back(0);
        m.ub.push_back(kInf);
        return k;
    };
    auto pairs = [&](const std::vector<std::string> &t, size_t first, auto apply) {
        if (t.size() <= first || (t.size() - first) % 2) throw std::runtime_error("Expected row/value pairs");
        for (size_t k = first; k < t.size(); k += 2) apply(t[k], number(t[k + 1]));
    };
    std::istringstream in(text);
    std::string line;
    while (std::getline(in, line)) {
        ++lineno;
        if (!line.empty() && line.back() == '\r') line.pop_back();
        if (trim(line).empty() || trim(line)[0] == '*') continue;
        auto t = words(line);
        // Fixed MPS allows blanks inside eight-character names (src/io.cpp:328-363).
        if (section == "ROWS" && line.size() >= 5 && line[0] == ' ' && trim(line.substr(2, 2)).empty() &&
            (line.size() <= 12 || trim(line.substr(12)).empty())) {
            auto name = trim(line.substr(4, 8));
            if (!name.empty()) t = {trim(line.substr(1, 1)), name};
        }
        if ((section == "COLUMNS" || section == "RHS" || section == "RANGES") && line.size() >= 36 && line[0] == ' ' &&
            trim(line.substr(12, 2)).empty() && trim(line.substr(22, 2)).empty() && trim(line.substr(36, 3)).empty()) {
            auto first_value = trim(line.substr(24, 12));
            auto first_row = trim(line.substr(14, 8));
            bool numeric = true;
            try {
                (void)number(first_value);
            } catch (const std::exception &) {
                numeric = false;
            }
            if (numeric && !first_row.empty()) {
                t = {trim(line.substr(4, 8)), first_row, first_value};
                if (line.size() > 39) {
                    auto second_row = trim(line.substr(39, 8));
                    if (!second_row.empty()) {
                        t.push_back(second_row);
                        t.push_back(line.size() > 49 ? trim(line.substr(49, 12)) : "");
                    }
                }
                if (t[0].empty()) t.erase(t.begin());
            }
        }
        if (t.empty()) continue;
        try {
            const std::string h = upper(t[0]);
            const bool header = line[0] != ' ' && line[0] != '\t';
            if (header) {
                if (h == "NAME") {
                    if (t.size() > 1) m.name = t[1];
                    section = h;
                    continue;
                }
                if (h == "ENDATA") {
                    ended = true;
                    break;
                }
                if (h == "OBJSENSE") {
                    section = h;
                    if (t.size() == 2) m.sense = sense_of(t[1]);
                    continue;
                }
                if (h == "QMATRIX" || h == "QUADOBJ")
                    throw std::runtime_error("UNSUPPORTED: quadratic objectives are outside this LP engine");
                if (h == "ROWS" || h == "COLUMNS" || h == "RHS" || h == "RANGES" || h == "BOUNDS") {
                    section = h;
                    continue;
                }
                throw std::runtime_error("UNSUPPORTED MPS section: " + h);
            }
            if (section == "OBJSENSE") {
                if (t.size() != 1) throw std::runtime_error("Malformed OBJSENSE");
                m.sense = sense_of(t[0]);
            } else if (section == "ROWS") {
                if (t.size() != 2 || t[0].size() != 1) throw std::runtime_error("Malformed ROWS");
                const char type = t[0][0];
                if (type == 'N') {
                    if (!obj.empty()) throw std::runtime_error("UNSUPPORTED multiple objective/free N rows");
                    obj = t[1];
                } else {
                    if (type != 'L' && type != 'G' && type != 'E') throw std::runtime_error("Unknown row type");
                    if (rows.count(t[1])) throw std::runtime_error("Duplicate row");
                    rows.emplace(t[1], int64_t(rows.size()));
                    m.row_names.push_back(t[1]);
                    rowtype.push_back(type);
                    rhs.push_back(0);
                    ranges.push_back(0);
                    has_range.push_back(false);
                }
            } else if (section == "COLUMNS") {
                if (t.size() == 3 && t[1] == "'MARKER'")
                    throw std::runtime_error("UNSUPPORTED: integer variables are outside this LP engine");
                size_t first = 1;
                std::string col = t[0];
                if (t.size() % 2 == 0) {
                    col = last_col;
                    first = 0;
                }
                if (col.empty()) throw std::runtime_error("Missing column name");
                last_col = col;
                const int64_t j = var(col);
                pairs(t, first, [&](const std::string &r, double v) {
                    if (r == obj)
                        m.c[j] += v;
                    else {
                        auto it = rows.find(r);
                        if (it == rows.end()) throw std::runtime_error("Unknown row: " + r);
                        e.push_back({it->second, j, v});
                    }
                });
            } else if (section == "RHS" || section == "RANGES") {
                auto &selected = section == "RHS" ? rhs_set : range_set;
                const size_t first = t.size() % 2 ? 1 : 0;
                if (first) {
                    if (selected.empty()) selected = t[0];
                    if (selected != t[0]) throw std::runtime_error("UNSUPPORTED multiple RHS/range sets");
                }
                pairs(t, first, [&](const std::string &r, double v) {
                    if (r == obj && section == "RHS") {
                        m.offset = -v;
                        return;
                    }
                    auto it = rows.find(r);
                    if (it == rows.end()) throw std::runtime_error("Unknown RHS row");
                    if (section == "RHS")
                        rhs[it->second] = v;
                    else {
                        ranges[it->second] = v;
                        has_range[it->second] = true;
                    }
                });
            } else if (section == "BOUNDS") {
                if (t.size() < 4 && line.size() >= 22 && trim(line.substr(3, 1)).empty() && trim(line.substr(12, 2)).empty() &&
                    trim(line.substr(22, 2)).empty()) {
                    auto set = trim(line.substr(4, 8));
                    auto name = trim(line.substr(14, 8));
                    auto value = line.size() > 24 ? trim(line.substr(24, 12)) : "";
                    if (!name.empty()) {
                        if (set.empty()) set = bound_set.empty() ? "DEFAULT" : bound_set;
                        t = {trim(line.substr(1, 2)), set, name};
                        if (!value.empty()) t.push_back(value);
                    }
                }
                if (t.size() < 3 || t.size() > 4) throw std::runtime_error("Malformed BOUNDS");
                if (bound_set.empty()) bound_set = t[1];
                if (bound_set != t[1]) throw std::runtime_error("UNSUPPORTED multiple bound sets");
                const auto type = upper(t[0]);
                if (type == "BV" || type == "LI" || type == "UI")
                    throw std::runtime_error("UNSUPPORTED: integer variables are outside this LP engine");
                const int64_t k = var(t[2]);
                if (!bound_seen.insert({type, t[2]}).second) throw std::runtime_error("Duplicate bound declaration");
                const bool value = type == "LO" || type == "UP" || type == "FX";
                if (t.size() != (value ? 4u : 3u)) throw std::runtime_error("Wrong bound field count");
                const double v = value ? number(t[3]) : 0;
                if (type == "LO")
                    m.lb[k] = v;
                else if (type == "UP") {
                    m.ub[k] = v;
                    if (v < 0 && !bound_seen.count({"LO", t[2]})) m.lb[k] = -kInf;
                } else if (type == "FX")
                    m.lb[k] = m.ub[k] = v;
                else if (type == "FR") {
                    m.lb[k] = -kInf;
                    m.ub[k] = kInf;
                } else if (type == "MI")
                    m.lb[k] = -kInf;
                else if (type == "PL")
                    m.ub[k] = kInf;
                else
                    throw std::runtime_error("UNSUPPORTED bound type: " + type);
            } else
                throw std::runtime_error("Unexpected MPS data");
        } catch (const std::exception &ex) {
            throw std::runtime_error("MPS line " + std::to_string(lineno) + ": " + ex.what());
        }
    }
    if (!ended || obj.empty()) throw std::runtime_error("Incomplete MPS / unclosed integer marker");
    for (size_t i = 0; i < rhs.size(); ++i) {
        double l = -kInf, u = kInf;
        if (rowtype[i] == 'E') l = u = rhs[i];
        if (rowtype[i] == 'L') u = rhs[i];
        if (rowtype[i] == 'G') l = rhs[i];
        if (has_range[i]) {
            const double r = ranges[i];
            if (rowtype[i] == 'L')
                l = u - std::abs(r);
            else if (rowtype[i] == 'G')
                u = l + std::abs(r);
            else if (r >= 0)
                u = l + r;
            else
                l = u + r;
        }
        m.rl.push_back(l);
        m.ru.push_back(u);
    }
    m.A = Csr::from_entries(int64_t(rows.size()), int64_t(vars.size()), std::move(e));
    canonicalize(m);
    return m;
}

namespace {

std::string number17(double v) {
    char buf[40];
    std::snprintf(buf, sizeof buf, "%.17g", v);
    return buf;
}

// Native JSON document of a model (src/io.cpp:765-798).  The reference's JSON
// objects keep their keys in sorted order, so the keys are written sorted here
// too (a constraint's coefficients by variable name); infinite bounds become
// null, which the reader maps back to the infinite default.
std::string json_document(const Model &m) {
    const int64_t n = m.cols(), rows = m.rows();
    std::vector<double> c(m.c);
    for (auto &v : c) v *= m.sense;
    Json doc = Json::object();
    Json cons = Json::array();
    std::vector<std::pair<std::string, double>> coef;
    for (int64_t i = 0; i < rows; ++i) {
        coef.clear();
        for (int64_t k = m.A.ptr[i]; k < m.A.ptr[i + 1]; ++k) coef.emplace_back(m.column_name(m.A.index[k]), m.A.value[k]);
        std::sort(coef.begin(), coef.end(), [](const auto &a, const auto &b) { return a.first < b.first; });
        Json row = Json::object();
        Json a = Json::object();
        for (const auto &[name, v] : coef) a[name] = v;
        row["coefficients"] = std::move(a);
        row["lb"] = m.rl[i];
        row["name"] = m.row_name(i);
        row["ub"] = m.ru[i];
        cons.push_back(std::move(row));
    }
    doc["constraints"] = std::move(cons);
    doc["name"] = m.name;
    Json obj = Json::object();
    obj["linear"] = Json(c);
    obj["offset"] = m.offset * m.sense;
    obj["quadratic_diagonal"] = Json(std::vector<double>(size_t(n), 0.0));
    doc["objective"] = std::move(obj);
    doc["sense"] = m.sense == 1 ? "min" : "max";
    Json vars = Json::array();
    for (int64_t j = 0; j < n; ++j) {
        Json v = Json::object();
        v["lb"] = m.lb[j];
        v["name"] = m.column_name(j);
        v["type"] = "continuous";
        v["ub"] = m.ub[j];
        vars.push_back(std::move(v));
    }
    doc["variables"] = std::move(vars);
    return doc.dump(2) + "\n";
}

// Free-format MPS in the layout of the reference's writer (src/io.cpp:815-901):
// OBJSENSE section, one RHS entry per row, RANGES for two-sided rows, FX/LO/MI/UP
// bounds, numbers at 17 significant digits.
std::string mps_document(const Model &m) {
    const int64_t n = m.cols(), rows = m.rows();
    for (int64_t j = 0; j < n; ++j)
        if (m.column_name(j).find_first_of(" \t\r\n") != std::string::npos)
            throw std::runtime_error("MPS export with whitespace in names requires fixed-format output; use JSON");
    for (int64_t i = 0; i < rows; ++i)
        if (m.row_name(i).find_first_of(" \t\r\n") != std::string::npos)
            throw std::runtime_error("MPS export with whitespace in names requires fixed-format output; use JSON");
    std::string out = "NAME          " + m.name + "\nOBJSENSE\n " + (m.sense == 1 ? "MIN" : "MAX") + "\nROWS\n N  OBJ\n";
    std::vector<double> rhs(static_cast<size_t>(rows));
    for (int64_t i = 0; i < rows; ++i) {
        const bool lo = std::isfinite(m.rl[i]), hi = std::isfinite(m.ru[i]);
        if (!lo && !hi) throw std::runtime_error("Cannot serialize fully free row to supported MPS subset");
        const char t = m.rl[i] == m.ru[i] ? 'E' : hi ? 'L' : 'G';
        rhs[i] = t == 'G' ? m.rl[i] : m.ru[i];
        out += std::string(" ") + t + "  " + m.row_name(i) + "\n";
    }
    out += "COLUMNS\n";
    const Csr at = m.A.transposed();
    for (int64_t j = 0; j < n; ++j) {
        const std::string name = m.column_name(j);
        out += " " + name + " OBJ " + number17(m.c[j] * m.sense) + "\n";
        for (int64_t k = at.ptr[j]; k < at.ptr[j + 1]; ++k)
            out += " " + name + " " + m.row_name(at.index[k]) + " " + number17(at.value[k]) + "\n";
    }
    out += "RHS\n RHS1 OBJ " + number17(-m.offset * m.sense) + "\n";
    for (int64_t i = 0; i < rows; ++i) out += " RHS1 " + m.row_name(i) + " " + number17(rhs[i]) + "\n";
    bool ranges = false;
    for (int64_t i = 0; i < rows; ++i)
        if (std::isfinite(m.rl[i]) && std::isfinite(m.ru[i]) && m.rl[i] != m.ru[i]) {
            if (!ranges) out += "RANGES\n";
            ranges = true;
            out += " RNG1 " + m.row_name(i) + " " + number17(m.ru[i] - m.rl[i]) + "\n";
        }
    out += "BOUNDS\n";
    for (int64_t j = 0; j < n; ++j) {
        const std::string name = m.column_name(j);
        if (m.lb[j] == m.ub[j]) {
            out += " FX BND " + name + " " + number17(m.lb[j]) + "\n";
            continue;
        }
        out += std::isfinite(m.lb[j]) ? " LO BND " + name + " " + number17(m.lb[j]) + "\n" : " MI BND " + name + "\n";
        if (std::isfinite(m.ub[j])) out += " UP BND " + name + " " + number17(m.ub[j]) + "\n";
    }
    return out + "ENDATA\n";
}

}  // namespace

void write_model(const Model &m, const std::string &path) {
    m.validate();
    const auto dot = path.find_last_of('.');
    const std::string ext = dot == std::string::npos ? "" : upper(path.substr(dot + 1));
    if (ext != "JSON" && ext != "MPS") throw std::runtime_error("Convert output must be .json or .mps");
    const std::string text = ext == "JSON" ? json_document(m) : mps_document(m);
    std::ofstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("Cannot write " + path);
    f << text;
    if (!f) throw std::runtime_error("Cannot write " + path);
}

Model read_model(const std::string &path) {
    auto dot = path.find_last_of('.');
    const std::string ext = dot == std::string::npos ? "" : upper(path.substr(dot + 1));
    if (ext == "JSON") return read_json_model(slurp(path));
    if (ext == "MPS") return read_mps_model(slurp(path));
    if (ext == "QPS" || ext == "QPLIB" || ext == "LP")
        throw std::runtime_error("UNSUPPORTED model format for this LP engine: ." + ext + " (use .mps or .json)");
    throw std::runtime_error("Supported formats: .mps .json");
}

Json result_json(const Model &m, const Result &r) {
    const auto &a = r.accuracy;
    std::vector<std::string> names(size_t(m.cols())), row_names(size_t(m.rows()));
    for (int64_t j = 0; j < m.cols(); ++j) names[j] = m.column_name(j);
    for (int64_t i = 0; i < m.rows(); ++i) row_names[i] = m.row_name(i);
    Json j = Json::object();
    j["solver"] = "NIRYUKTI";
    j["version"] = engine_version();
    j["status"] = r.status;
    j["message"] = r.message;
    j["problem_type"] = "LP";
    Json model = Json::object();
    model["name"] = m.name;
    model["variable_names"] = Json(names);
    model["row_names"] = Json(row_names);
    model["fingerprint"] = m.fingerprint();
    model["rows"] = (long long)m.rows();
    model["columns"] = (long long)m.cols();
    model["nonzeros"] = (long long)m.A.nnz();
    j["model"] = std::move(model);
    j["objective"] = a.finite ? Json(m.sense * a.objective) : Json();
    Json acc = Json::object();
    acc["finite"] = a.finite;
    acc["primal_residual"] = a.primal;
    acc["dual_residual"] = a.dual;
    acc["relative_gap"] = a.gap;
    acc["kkt_error"] = a.kkt;
    acc["integrality"] = a.finite ? 0.0 : kInf;
    acc["primal_absolute"] = a.primal_absolute;
    acc["dual_absolute"] = a.dual_absolute;
    acc["complementarity"] = a.complementarity;
    acc["dual_bound"] = a.lower_bound;
    j["accuracy"] = std::move(acc);
    Json perf = Json::object();
    perf["total_seconds"] = r.seconds;
    perf["preprocess_seconds"] = r.preprocess_seconds;
    perf["transfer_setup_seconds"] = r.transfer_seconds;
    perf["iteration_seconds"] = r.iteration_seconds;
    perf["verification_seconds"] = r.verification_seconds;
    perf["iterations"] = (long long)r.iterations;
    perf["weight_updates"] = (long long)r.weight_updates;
    perf["polishing_iterations"] = (long long)r.polishing_iterations;
    perf["polishing_attempts"] = (long long)r.polishing_attempts;
    perf["operator_norm_estimate"] = r.operator_norm_estimate;
    perf["restarts"] = (long long)r.restarts;
    perf["rejected_steps"] = (long long)r.rejected_steps;
    perf["monitor_checks"] = (long long)r.device_candidate_checks;
    perf["host_candidate_checks"] = (long long)r.host_candidate_checks;
    perf["skipped_candidate_checks"] = 0;
    j["performance"] = std::move(perf);
    Json hw = Json::object();
    hw["backend"] = r.backend;
    hw["device"] = r.device_name;
    hw["precision"] = r.matrix_precision == "mixed" ? "FP32 matrix values, FP64 iterates"
                      : r.matrix_precision == "fp32" ? "FP32 SpMV arithmetic, FP64 iterates, FP64 refinement"
                                                     : "FP64";
    j["hardware"] = std::move(hw);
    Json pre = Json::object();
    pre["removed_rows"] = (long long)r.removed_rows;
    pre["removed_columns"] = (long long)r.removed_columns;
    pre["bounds_tightened"] = 0;
    j["presolve"] = std::move(pre);
    j["primal"] = Json(r.x);
    j["dual"] = Json(r.y);
    Json sel = Json::object();
    sel["automatic"] = r.automatic;
    sel["requested_method"] = r.method_requested;
    sel["method"] = r.method_selected;
    sel["relaxation_method"] = "";
    sel["device"] = r.backend;
    sel["reason"] = r.device_reason;
    sel["estimated_gpu_bytes"] = r.estimated_gpu_bytes;
    Json cfg = Json::object();
    cfg["iteration_limit"] = (long long)r.iteration_limit;
    cfg["precision"] = r.matrix_precision;
    cfg["scaling"] = r.scaling;
    cfg["scaling_passes"] = r.scaling_passes;
    cfg["time_limit_seconds"] = r.time_limit;
    cfg["tolerance"] = r.tol;
    sel["configuration"] = std::move(cfg);
    j["selection"] = std::move(sel);
    Json gx = Json::object();
    gx["index_bits"] = r.gpu_index_bits;
    gx["cuda_graphs"] = r.graph_execution;
    gx["matrix_precision"] = r.matrix_precision;
    if (r.matrix_precision == "fp32") gx["fp32_iterations"] = (long long)r.fp32_iterations;
    gx["device_prepare"] = r.device_prepare;
    j["gpu_execution"] = std::move(gx);
    if (!r.infeasibility_ray.empty()) {
        Json cert = Json::object();
        cert["kind"] = "BOX_ROW_FARKAS";
        cert["dual_ray"] = Json(r.infeasibility_ray);
        cert["verified_margin"] = r.certificate_margin;
        j["certificate"] = std::move(cert);
    }
    return j;
}

Solution read_solution(const Model &m, const std::string &path, bool allow_model_change) {
    std::ifstream f(path, std::ios::binary | std::ios::ate);
    if (!f || f.tellg() > 64 * 1024 * 1024) throw std::runtime_error("Invalid or oversized solution file");
    f.seekg(0);
    std::string raw((std::istreambuf_iterator<char>(f)), {});
    Json j = Json::parse(raw, 32);
    const Json &model = j.at("model");
    if (model.at("fingerprint").str() != m.fingerprint()) {
        if (!allow_model_change)
            throw std::runtime_error("Solution/model fingerprint mismatch; parametric solves require --allow-model-change");
        const Json &vn = model.at("variable_names"), &rn = model.at("row_names");
        bool same = vn.size() == size_t(m.cols()) && rn.size() == size_t(m.rows());
        for (int64_t k = 0; same && k < m.cols(); ++k) same = vn[k].str() == m.column_name(k);
        for (int64_t k = 0; same && k < m.rows(); ++k) same = rn[k].str() == m.row_name(k);
        if (!same) throw std::runtime_error("Warm-start row/variable names or ordering mismatch");
    }
    Solution s;
    s.x = j.at("primal").numbers();
    s.y = j.at("dual").numbers();
    s.status = j.at("status").str();
    if (const Json *o = j.find("objective")) {
        if (o->is_number())
            s.objective = o->number();
        else if (s.status == "OPTIMAL")
            throw std::runtime_error("Claimed optimal objective must be finite and numeric");
    }
    if (const Json *c = j.find("certificate")) {
        if (c->at("kind").str() != "BOX_ROW_FARKAS") throw std::runtime_error("Unsupported certificate kind");
        s.infeasibility_ray = c->at("dual_ray").numbers();
    }
    return s;
}

}  // namespace nyk
