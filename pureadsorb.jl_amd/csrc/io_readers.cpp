// io_readers.cpp -- the input files of the Widom path read in C++ for hosts without Julia:
// the minimal P1 CIF reader (src/structure.jl:37-69), the kUPS-style force-field YAML
// (src/forcefield.jl:54-61) and guest YAML (src/forcefield.jl:91-101). Host code only; no
// device is touched. Errors keep the reference's ArgumentError wording.
//
// The YAML side reads the subset those files use: block mappings, block sequences (also at the
// indentation of their key), flow sequences, plain / single / double quoted scalars, nulls
// (`~`, `null`, empty) and `#` comments. Anything else (flow mappings, block scalars, anchors)
// is reported as unsupported instead of guessed.
#include <pureadsorb_amd.h>

#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <memory>
#include <sstream>
#include <string>
#include <utility>
#include <vector>

namespace pa {
int fail(int code, const char *fmt, ...);
int ok();
}  // namespace pa

struct pa_cif_file {
    double cell[9];
    std::vector<double> frac, charges;
    std::vector<std::string> labels, symbols;
};

struct pa_ff_file {
    std::vector<std::string> names;
    std::vector<double> sigma, epsilon;  // per type, unmixed
    double cutoff = 0.0;
    bool tail = true;
};

struct pa_guest_file {
    std::vector<double> sites, charges;
    std::vector<int32_t> types;
    double tc = 1.0, pc = 1.0, omega = 0.0;
};

namespace {

using pa::fail;

const char *WS = " \t\r\n\f\v";

std::string strip(const std::string &s, const char *chars = WS)
{
    const size_t a = s.find_first_not_of(chars);
    if (a == std::string::npos) return std::string();
    const size_t b = s.find_last_not_of(chars);
    return s.substr(a, b - a + 1);
}

bool starts_with(const std::string &s, const std::string &p) { return s.compare(0, p.size(), p) == 0; }

std::vector<std::string> split_ws(const std::string &s)
{
    std::vector<std::string> out;
    std::istringstream is(s);
    std::string t;
    while (is >> t) out.push_back(t);
    return out;
}

// Julia's parse(Float64, s): the whole (whitespace-stripped) string must be one number.
bool parse_f64(const std::string &s0, double *out)
{
    const std::string s = strip(s0);
    if (s.empty()) return false;
    errno = 0;
    char *end = nullptr;
    const double v = std::strtod(s.c_str(), &end);
    if (end != s.c_str() + s.size()) return false;
    *out = v;
    return true;
}

int read_lines(const char *path, std::vector<std::string> *lines)
{
    if (!path) return fail(PA_ERR_ARG, "NULL path");
    std::ifstream in(path);
    if (!in) return fail(PA_ERR_ARG, "cannot open %s: %s", path, std::strerror(errno));
    std::string ln;
    while (std::getline(in, ln)) lines->push_back(ln);
    return PA_OK;
}

// sin and cos of an angle in degrees, exactly 0 / +-1 at multiples of 90 (Julia sind / cosd).
void sincosd(double deg, double *s, double *c)
{
    const double n = std::nearbyint(deg / 90.0);
    const double r = (deg - 90.0 * n) * (M_PI / 180.0);
    const double sr = std::sin(r), cr = std::cos(r);
    const long q = ((long)n % 4 + 4) % 4;
    switch (q) {
    case 0: *s = sr; *c = cr; break;
    case 1: *s = cr; *c = 0.0 - sr; break;
    case 2: *s = 0.0 - sr; *c = 0.0 - cr; break;
    default: *s = 0.0 - cr; *c = sr; break;
    }
}

// src/cell.jl:2-12: lattice vectors as columns, a along x, b in the xy plane (column-major).
int cell_matrix(const double p[6], double A[9])
{
    const double a = p[0], b = p[1], c = p[2];
    if (!(a > 0 && b > 0 && c > 0))
        return fail(PA_ERR_ARG, "cell lengths must be positive, got a=%g, b=%g, c=%g", a, b, c);
    double sa, ca, sb, cb, sg, cg;
    sincosd(p[3], &sa, &ca);
    sincosd(p[4], &sb, &cb);
    sincosd(p[5], &sg, &cg);
    if (sg == 0.0) return fail(PA_ERR_ARG, "cell angle γ=%g is degenerate: sin(γ) is zero", p[5]);
    const double cy = (ca - cb * cg) / sg;
    const double rad = 1.0 - cb * cb - cy * cy;
    if (!(rad > 0))
        return fail(PA_ERR_ARG, "cell angles α=%g, β=%g, γ=%g describe a degenerate cell", p[3], p[4], p[5]);
    const double cz = std::sqrt(rad);
    const double M[9] = {a, 0.0, 0.0, b * cg, b * sg, 0.0, c * cb, c * cy, c * cz};
    std::memcpy(A, M, sizeof(M));
    return PA_OK;
}

int copy_strings(const std::vector<std::string> &v, char *buf, size_t cap, size_t *needed)
{
    size_t n = 0;
    for (const std::string &s : v) n += s.size() + 1;
    if (needed) *needed = n;
    if (!buf) return pa::ok();
    if (cap < n) return fail(PA_ERR_DIM, "string buffer holds %zu bytes, %zu needed", cap, n);
    for (const std::string &s : v) {
        std::memcpy(buf, s.c_str(), s.size() + 1);
        buf += s.size() + 1;
    }
    return pa::ok();
}

// ---------------------------------------------------------------------------------------------
// YAML subset
// ---------------------------------------------------------------------------------------------
struct YNode {
    enum Kind { NUL, SCALAR, SEQ, MAP } kind = NUL;
    std::string text;  // SCALAR
    bool quoted = false;
    std::vector<YNode> items;                        // SEQ
    std::vector<std::pair<std::string, YNode>> map;  // MAP, file order

    const YNode *get(const std::string &k) const
    {
        const YNode *r = nullptr;
        for (const auto &kv : map)
            if (kv.first == k) r = &kv.second;  // a repeated key keeps its last value
        return r;
    }
};

struct YLine {
    int indent;
    std::string text;  // without indentation and comment
    int lineno;
};

struct YParser {
    std::vector<YLine> L;
    size_t i = 0;
    std::string err;

    bool error(int lineno, const std::string &what)
    {
        if (err.empty()) err = "line " + std::to_string(lineno) + ": " + what;
        return false;
    }

    static std::string drop_comment(const std::string &s)
    {
        char q = 0;
        for (size_t k = 0; k < s.size(); ++k) {
            const char ch = s[k];
            if (q) {
                if (ch == q) q = 0;
            } else if (ch == '\'' || ch == '"') {
                q = ch;
            } else if (ch == '#' && (k == 0 || s[k - 1] == ' ' || s[k - 1] == '\t')) {
                return s.substr(0, k);
            }
        }
        return s;
    }

    void load(const std::vector<std::string> &raw)
    {
        for (size_t k = 0; k < raw.size(); ++k) {
            std::string s = drop_comment(raw[k]);
            const std::string t = strip(s);
            if (t.empty() || t == "---" || t == "...") continue;
            int ind = 0;
            while (ind < (int)s.size() && s[ind] == ' ') ++ind;
            L.push_back({ind, strip(s), (int)k + 1});
        }
    }

    static bool is_null_text(const std::string &t) { return t.empty() || t == "~" || t == "null" || t == "Null" || t == "NULL"; }

    bool scalar(const std::string &t0, int lineno, YNode *out)
    {
        const std::string t = strip(t0);
        if (is_null_text(t)) {
            out->kind = YNode::NUL;
            return true;
        }
        if (t[0] == '{' || t[0] == '|' || t[0] == '>' || t[0] == '&' || t[0] == '*' || t[0] == '!')
            return error(lineno, "unsupported YAML construct '" + t + "'");
        out->kind = YNode::SCALAR;
        if (t[0] == '\'' || t[0] == '"') {
            const char q = t[0];
            if (t.size() < 2 || t.back() != q) return error(lineno, "unterminated quoted scalar " + t);
            std::string body = t.substr(1, t.size() - 2);
            if (q == '\'') {  // '' is an escaped quote
                std::string r;
                for (size_t k = 0; k < body.size(); ++k) {
                    r += body[k];
                    if (body[k] == '\'' && k + 1 < body.size() && body[k + 1] == '\'') ++k;
                }
                body = r;
            } else if (body.find('\\') != std::string::npos) {
                return error(lineno, "escape sequences in double-quoted scalars are not supported");
            }
            out->text = body;
            out->quoted = true;
        } else {
            out->text = t;
        }
        return true;
    }

    // flow sequence starting at s[pos] == '['; pos ends after the matching ']'
    bool flow_seq(const std::string &s, size_t &pos, int lineno, YNode *out)
    {
        out->kind = YNode::SEQ;
        ++pos;
        for (;;) {
            while (pos < s.size() && (s[pos] == ' ' || s[pos] == '\t')) ++pos;
            if (pos >= s.size()) return error(lineno, "unterminated flow sequence");
            if (s[pos] == ']') {
                ++pos;
                return true;
            }
            YNode item;
            if (s[pos] == '[') {
                if (!flow_seq(s, pos, lineno, &item)) return false;
            } else {
                size_t e = pos;
                char q = 0;
                while (e < s.size()) {
                    const char ch = s[e];
                    if (q) {
                        if (ch == q) q = 0;
                    } else if (ch == '\'' || ch == '"') {
                        q = ch;
                    } else if (ch == ',' || ch == ']') {
                        break;
                    }
                    ++e;
                }
                if (!scalar(s.substr(pos, e - pos), lineno, &item)) return false;
                pos = e;
            }
            out->items.push_back(std::move(item));
            while (pos < s.size() && (s[pos] == ' ' || s[pos] == '\t')) ++pos;
            if (pos < s.size() && s[pos] == ',') ++pos;
            else if (pos < s.size() && s[pos] == ']') continue;
            else return error(lineno, "malformed flow sequence");
        }
    }

    // inline value after "key:" or "- "; a flow sequence may continue on following lines
    bool inline_value(std::string v, int lineno, YNode *out)
    {
        if (!v.empty() && v[0] == '[') {
            int depth = 0;
            auto balance = [&](const std::string &s) {
                char q = 0;
                for (char ch : s) {
                    if (q) { if (ch == q) q = 0; }
                    else if (ch == '\'' || ch == '"') q = ch;
                    else if (ch == '[') ++depth;
                    else if (ch == ']') --depth;
                }
            };
            balance(v);
            while (depth > 0 && i < L.size()) {
                v += " " + L[i].text;
                depth = 0;
                balance(v);
                ++i;
            }
            size_t pos = 0;
            if (!flow_seq(v, pos, lineno, out)) return false;
            if (!strip(v.substr(pos)).empty()) return error(lineno, "trailing text after flow sequence");
            return true;
        }
        return scalar(v, lineno, out);
    }

    static bool is_seq_item(const std::string &t) { return t == "-" || starts_with(t, "- "); }

    // position of the ':' ending a mapping key (followed by a space or the end), outside quotes
    static size_t key_colon(const std::string &t)
    {
        char q = 0;
        for (size_t k = 0; k < t.size(); ++k) {
            const char ch = t[k];
            if (q) {
                if (ch == q) q = 0;
            } else if (ch == '\'' || ch == '"') {
                q = ch;
            } else if (ch == ':' && (k + 1 == t.size() || t[k + 1] == ' ' || t[k + 1] == '\t')) {
                return k;
            }
        }
        return std::string::npos;
    }

    bool block(int ind, YNode *out)
    {
        if (is_seq_item(L[i].text)) return seq(ind, out);
        return mapping(ind, out);
    }

    bool seq(int ind, YNode *out)
    {
        out->kind = YNode::SEQ;
        while (i < L.size() && L[i].indent == ind && is_seq_item(L[i].text)) {
            const int lineno = L[i].lineno;
            const std::string rest = strip(L[i].text.substr(1));
            ++i;
            YNode item;
            if (rest.empty()) {
                if (i < L.size() && L[i].indent > ind) {
                    if (!block(L[i].indent, &item)) return false;
                }
            } else {
                if (key_colon(rest) != std::string::npos && rest[0] != '[' && rest[0] != '\'' && rest[0] != '"')
                    return error(lineno, "mappings inside sequence items are not supported");
                if (!inline_value(rest, lineno, &item)) return false;
            }
            out->items.push_back(std::move(item));
        }
        if (i < L.size() && L[i].indent > ind) return error(L[i].lineno, "unexpected indentation");
        return true;
    }

    bool mapping(int ind, YNode *out)
    {
        out->kind = YNode::MAP;
        while (i < L.size() && L[i].indent == ind && !is_seq_item(L[i].text)) {
            const int lineno = L[i].lineno;
            const std::string t = L[i].text;
            const size_t c = key_colon(t);
            if (c == std::string::npos) return error(lineno, "expected 'key: value', got '" + t + "'");
            YNode key;
            if (!scalar(t.substr(0, c), lineno, &key)) return false;
            const std::string k = key.kind == YNode::NUL ? std::string() : key.text;
            const std::string v = strip(t.substr(c + 1));
            ++i;
            YNode val;
            if (v.empty()) {
                if (i < L.size() && L[i].indent > ind) {
                    if (!block(L[i].indent, &val)) return false;
                } else if (i < L.size() && L[i].indent == ind && is_seq_item(L[i].text)) {
                    if (!seq(ind, &val)) return false;
                }
            } else if (!inline_value(v, lineno, &val)) {
                return false;
            }
            out->map.emplace_back(k, std::move(val));
        }
        if (i < L.size() && L[i].indent > ind) return error(L[i].lineno, "unexpected indentation");
        return true;
    }
};

int load_yaml(const char *path, YNode *root)
{
    std::vector<std::string> raw;
    int st = read_lines(path, &raw);
    if (st) return st;
    YParser P;
    P.load(raw);
    if (P.L.empty()) return fail(PA_ERR_ARG, "%s: empty YAML document", path);
    if (P.L[0].indent != 0 || YParser::is_seq_item(P.L[0].text))
        return fail(PA_ERR_ARG, "%s: the document must be a mapping", path);
    if (!P.mapping(0, root) || P.i != P.L.size())
        return fail(PA_ERR_ARG, "%s: %s", path, P.err.empty() ? "unexpected content" : P.err.c_str());
    return PA_OK;
}

// T(x) of a YAML number (src/forcefield.jl: `T(d[key])`); `underscores` also accepts the
// digit-group form that parse_number strips ("7_840_000").
int yaml_number(const YNode *n, const char *what, bool underscores, double *out)
{
    if (!n || n->kind == YNode::NUL) return fail(PA_ERR_ARG, "%s is missing", what);
    if (n->kind != YNode::SCALAR) return fail(PA_ERR_ARG, "%s must be a number", what);
    if (n->quoted && !underscores) return fail(PA_ERR_ARG, "%s must be a number, got the string \"%s\"", what, n->text.c_str());
    std::string t = n->text;
    if (underscores) {
        std::string r;
        for (char ch : t)
            if (ch != '_') r += ch;
        t = r;
    }
    if (!parse_f64(t, out)) return fail(PA_ERR_ARG, "cannot parse \"%s\" as Float64 (%s)", n->text.c_str(), what);
    return PA_OK;
}

int yaml_seq(const YNode *n, const char *what, const YNode **out)
{
    if (!n) return fail(PA_ERR_ARG, "%s is missing", what);
    if (n->kind != YNode::SEQ) return fail(PA_ERR_ARG, "%s must be a sequence", what);
    *out = n;
    return PA_OK;
}

}  // namespace

// ---------------------------------------------------------------------------------------------
// CIF (src/structure.jl:37-69)
// ---------------------------------------------------------------------------------------------
extern "C" int pa_read_cif(const char *path, pa_cif_file **out)
{
    if (!out) return fail(PA_ERR_ARG, "NULL argument");
    *out = nullptr;
    std::vector<std::string> raw;
    int st = read_lines(path, &raw);
    if (st) return st;
    std::vector<std::string> lines;
    lines.reserve(raw.size());
    for (const std::string &l : raw) lines.push_back(strip(l));
    std::string val;
    auto getval = [&](const std::string &key) -> int {
        for (const std::string &l : lines)
            if (starts_with(l, key + " ") || starts_with(l, key + "\t")) {
                val = strip(l.substr(key.size()));
                return PA_OK;
            }
        return fail(PA_ERR_ARG, "CIF is missing %s", key.c_str());
    };
    if ((st = getval("_symmetry_space_group_name_H-M"))) return st;
    const std::string sg = strip(val, "'\"");
    std::string compact;
    for (char ch : sg)
        if (ch != ' ') compact += ch;
    if (compact != "P1") return fail(PA_ERR_ARG, "only P1 CIF files are supported, got space group %s", sg.c_str());
    static const char *keys6[6] = {"_cell_length_a", "_cell_length_b", "_cell_length_c",
                                   "_cell_angle_alpha", "_cell_angle_beta", "_cell_angle_gamma"};
    double p[6];
    for (int k = 0; k < 6; ++k) {
        if ((st = getval(keys6[k]))) return st;
        if (!parse_f64(val, &p[k])) return fail(PA_ERR_ARG, "cannot parse \"%s\" as Float64", val.c_str());
    }
    std::unique_ptr<pa_cif_file> f(new pa_cif_file());
    if ((st = cell_matrix(p, f->cell))) return st;
    size_t i = 0;
    while (i < lines.size() && lines[i] != "_atom_site_label") ++i;
    if (i == lines.size()) return fail(PA_ERR_ARG, "CIF has no _atom_site_label loop");
    std::vector<std::string> cols;
    while (i < lines.size() && starts_with(lines[i], "_atom_site_")) cols.push_back(lines[i++]);
    auto col = [&](const char *name) -> int {
        for (size_t k = 0; k < cols.size(); ++k)
            if (cols[k] == name) return (int)k;
        return -1;
    };
    if (col("_atom_site_charge") < 0)
        return fail(PA_ERR_ARG, "CIF has no _atom_site_charge column; partial charges are required");
    const int cx = col("_atom_site_fract_x"), cy = col("_atom_site_fract_y"), cz = col("_atom_site_fract_z");
    const int cq = col("_atom_site_charge"), cl = col("_atom_site_label"), cs = col("_atom_site_type_symbol");
    if (cx < 0 || cy < 0 || cz < 0 || cl < 0 || cs < 0)
        return fail(PA_ERR_ARG, "CIF atom_site loop lacks label, type_symbol or fractional coordinates");
    while (i < lines.size() && !lines[i].empty() && !starts_with(lines[i], "_") && !starts_with(lines[i], "loop_")) {
        const std::vector<std::string> fld = split_ws(lines[i]);
        if (fld.size() != cols.size())
            return fail(PA_ERR_ARG, "CIF atom row has %zu fields, header has %zu", fld.size(), cols.size());
        double v[4];
        const int ci[4] = {cx, cy, cz, cq};
        for (int k = 0; k < 4; ++k)
            if (!parse_f64(fld[ci[k]], &v[k])) return fail(PA_ERR_ARG, "cannot parse \"%s\" as Float64", fld[ci[k]].c_str());
        f->frac.insert(f->frac.end(), v, v + 3);
        f->charges.push_back(v[3]);
        f->labels.push_back(fld[cl]);
        f->symbols.push_back(fld[cs]);
        ++i;
    }
    if (f->charges.size() > (size_t)INT32_MAX) return fail(PA_ERR_ARG, "CIF has too many atoms");
    *out = f.release();
    return pa::ok();
}

extern "C" int pa_cif_info(const pa_cif_file *f, int32_t *natoms, double cell[9])
{
    if (!f) return fail(PA_ERR_ARG, "NULL CIF handle");
    if (natoms) *natoms = (int32_t)f->charges.size();
    if (cell) std::memcpy(cell, f->cell, sizeof(f->cell));
    return pa::ok();
}

extern "C" int pa_cif_atoms(const pa_cif_file *f, double *frac, double *charges)
{
    if (!f) return fail(PA_ERR_ARG, "NULL CIF handle");
    if (frac && !f->frac.empty()) std::memcpy(frac, f->frac.data(), sizeof(double) * f->frac.size());
    if (charges && !f->charges.empty()) std::memcpy(charges, f->charges.data(), sizeof(double) * f->charges.size());
    return pa::ok();
}

extern "C" int pa_cif_strings(const pa_cif_file *f, int32_t which, char *buf, size_t cap, size_t *needed)
{
    if (!f) return fail(PA_ERR_ARG, "NULL CIF handle");
    if (which != PA_CIF_LABELS && which != PA_CIF_SYMBOLS) return fail(PA_ERR_ARG, "unknown string column %d", which);
    return copy_strings(which == PA_CIF_LABELS ? f->labels : f->symbols, buf, cap, needed);
}

extern "C" void pa_cif_destroy(pa_cif_file *f) { delete f; }

// ---------------------------------------------------------------------------------------------
// Force field (src/forcefield.jl:54-61) and Lorentz-Berthelot tables (:24-39)
// ---------------------------------------------------------------------------------------------
extern "C" int pa_read_forcefield(const char *path, pa_ff_file **out)
{
    if (!out) return fail(PA_ERR_ARG, "NULL argument");
    *out = nullptr;
    YNode d;
    int st = load_yaml(path, &d);
    if (st) return st;
    const YNode *params = d.get("parameters");
    if (!params || params->kind != YNode::MAP) return fail(PA_ERR_ARG, "%s: parameters must be a mapping of type names", path);
    std::unique_ptr<pa_ff_file> f(new pa_ff_file());
    for (const auto &kv : params->map) {
        bool seen = false;
        for (const std::string &n : f->names) seen = seen || n == kv.first;
        if (seen) continue;  // a repeated name keeps its last value (read below through get)
        const YNode *e = params->get(kv.first);
        if (e->kind != YNode::SEQ || e->items.size() < 2)
            return fail(PA_ERR_ARG, "%s: parameters of %s must be a [sigma, epsilon] sequence", path, kv.first.c_str());
        double sg = 1.0, ep = 0.0;  // missing sigma = 1, missing epsilon = 0
        const std::string ws = "sigma of " + kv.first, we = "epsilon of " + kv.first;
        if (e->items[0].kind != YNode::NUL && (st = yaml_number(&e->items[0], ws.c_str(), false, &sg))) return st;
        if (e->items[1].kind != YNode::NUL && (st = yaml_number(&e->items[1], we.c_str(), false, &ep))) return st;
        f->names.push_back(kv.first);
        f->sigma.push_back(sg);
        f->epsilon.push_back(ep);
    }
    if ((st = yaml_number(d.get("cutoff"), "cutoff", false, &f->cutoff))) return st;
    const YNode *tc = d.get("tail_correction");
    if (tc && tc->kind != YNode::NUL) {
        const std::string &t = tc->text;
        if (tc->kind == YNode::SCALAR && !tc->quoted && (t == "true" || t == "True" || t == "TRUE")) f->tail = true;
        else if (tc->kind == YNode::SCALAR && !tc->quoted && (t == "false" || t == "False" || t == "FALSE")) f->tail = false;
        else return fail(PA_ERR_ARG, "%s: tail_correction must be true or false", path);
    }
    *out = f.release();
    return pa::ok();
}

extern "C" int pa_ff_info(const pa_ff_file *f, int32_t *ntypes, double *cutoff, int32_t *tail)
{
    if (!f) return fail(PA_ERR_ARG, "NULL force-field handle");
    if (ntypes) *ntypes = (int32_t)f->names.size();
    if (cutoff) *cutoff = f->cutoff;
    if (tail) *tail = f->tail ? 1 : 0;
    return pa::ok();
}

extern "C" int pa_ff_params(const pa_ff_file *f, double *sigma_type, double *epsilon_type, double *sigma,
                            double *epsilon)
{
    if (!f) return fail(PA_ERR_ARG, "NULL force-field handle");
    const size_t n = f->names.size();
    for (size_t a = 0; a < n; ++a) {
        if (sigma_type) sigma_type[a] = f->sigma[a];
        if (epsilon_type) epsilon_type[a] = f->epsilon[a];
        for (size_t b = 0; b < n; ++b) {  // column-major [a + n b]
            if (sigma) sigma[a + n * b] = (f->sigma[a] + f->sigma[b]) / 2;
            if (epsilon) epsilon[a + n * b] = std::sqrt(f->epsilon[a] * f->epsilon[b]);
        }
    }
    return pa::ok();
}

extern "C" int pa_ff_names(const pa_ff_file *f, char *buf, size_t cap, size_t *needed)
{
    if (!f) return fail(PA_ERR_ARG, "NULL force-field handle");
    return copy_strings(f->names, buf, cap, needed);
}

namespace {
int typeindex(const pa_ff_file *f, const std::string &name, int32_t *index)
{
    for (size_t k = 0; k < f->names.size(); ++k)
        if (f->names[k] == name) {
            *index = (int32_t)k;
            return PA_OK;
        }
    std::string all = "[";
    for (size_t k = 0; k < f->names.size(); ++k) all += (k ? ", \"" : "\"") + f->names[k] + "\"";
    all += "]";
    return fail(PA_ERR_ARG, "unknown LJ type %s; force field defines %s", name.c_str(), all.c_str());
}
}  // namespace

extern "C" int pa_ff_typeindex(const pa_ff_file *f, const char *name, int32_t *index)
{
    if (!f || !name || !index) return fail(PA_ERR_ARG, "NULL argument");
    const int st = typeindex(f, name, index);
    return st ? st : pa::ok();
}

extern "C" int pa_cif_types(const pa_cif_file *c, const pa_ff_file *f, int32_t *types)
{
    if (!c || !f || !types) return fail(PA_ERR_ARG, "NULL argument");
    for (size_t k = 0; k < c->symbols.size(); ++k) {
        const int st = typeindex(f, c->symbols[k] + "_", &types[k]);
        if (st) return st;
    }
    return pa::ok();
}

extern "C" void pa_ff_destroy(pa_ff_file *f) { delete f; }

// ---------------------------------------------------------------------------------------------
// Guest (src/forcefield.jl:91-101)
// ---------------------------------------------------------------------------------------------
extern "C" int pa_read_guest(const char *path, const pa_ff_file *ff, pa_guest_file **out)
{
    if (!out || !ff) return fail(PA_ERR_ARG, "NULL argument");
    *out = nullptr;
    YNode d;
    int st = load_yaml(path, &d);
    if (st) return st;
    const YNode *pos, *sym, *chg;
    if ((st = yaml_seq(d.get("positions"), "positions", &pos))) return st;
    if ((st = yaml_seq(d.get("symbols"), "symbols", &sym))) return st;
    if ((st = yaml_seq(d.get("charges"), "charges", &chg))) return st;
    const size_t n = pos->items.size();
    if (sym->items.size() != n || chg->items.size() != n)
        return fail(PA_ERR_DIM, "guest has %zu positions, %zu symbols and %zu charges", n, sym->items.size(),
                    chg->items.size());
    std::unique_ptr<pa_guest_file> g(new pa_guest_file());
    for (size_t s = 0; s < n; ++s) {
        const YNode &p = pos->items[s];
        if (p.kind != YNode::SEQ || p.items.size() != 3)
            return fail(PA_ERR_DIM, "guest position %zu must have 3 coordinates", s + 1);
        for (int d3 = 0; d3 < 3; ++d3) {
            double v;
            if ((st = yaml_number(&p.items[d3], "guest position", false, &v))) return st;
            g->sites.push_back(v);
        }
        const YNode &y = sym->items[s];
        if (y.kind != YNode::SCALAR) return fail(PA_ERR_ARG, "guest symbol %zu must be a string", s + 1);
        int32_t t;
        if ((st = typeindex(ff, y.text, &t))) return st;
        g->types.push_back(t);
        double q;
        if ((st = yaml_number(&chg->items[s], "guest charge", false, &q))) return st;
        g->charges.push_back(q);
    }
    if ((st = yaml_number(d.get("critical_temperature"), "critical_temperature", false, &g->tc))) return st;
    if ((st = yaml_number(d.get("critical_pressure"), "critical_pressure", true, &g->pc))) return st;
    if ((st = yaml_number(d.get("acentric_factor"), "acentric_factor", false, &g->omega))) return st;
    *out = g.release();
    return pa::ok();
}

extern "C" int pa_guest_info(const pa_guest_file *g, int32_t *nsites, double *tc, double *pc, double *omega)
{
    if (!g) return fail(PA_ERR_ARG, "NULL guest handle");
    if (nsites) *nsites = (int32_t)g->types.size();
    if (tc) *tc = g->tc;
    if (pc) *pc = g->pc;
    if (omega) *omega = g->omega;
    return pa::ok();
}

extern "C" int pa_guest_sites(const pa_guest_file *g, double *sites, int32_t *types, double *charges)
{
    if (!g) return fail(PA_ERR_ARG, "NULL guest handle");
    if (sites && !g->sites.empty()) std::memcpy(sites, g->sites.data(), sizeof(double) * g->sites.size());
    if (types && !g->types.empty()) std::memcpy(types, g->types.data(), sizeof(int32_t) * g->types.size());
    if (charges && !g->charges.empty()) std::memcpy(charges, g->charges.data(), sizeof(double) * g->charges.size());
    return pa::ok();
}

extern "C" void pa_guest_destroy(pa_guest_file *g) { delete g; }
