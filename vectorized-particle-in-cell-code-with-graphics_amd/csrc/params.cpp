// params.cpp -- pic_params defaults and the parameters.h loader.
//
// The reference configures itself at compile time through the #define lines of
// parameters.h and constants.h (parameters.h:41-366, constants.h:15-36).  This
// file evaluates such a file at run time: a small C-expression evaluator with
// C's int/double arithmetic rules, casts, string literals and the math
// functions the reference's macros use.  Macros are evaluated lazily, so the
// display/diagnostic macros this library does not use never need to parse.
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <fstream>
#include <map>
#include <sstream>
#include <string>

#include "../../include/pic2d.h"
#include "pic_host.h"

namespace {

// constants.h:15-36 (names and values as the reference defines them)
struct ConstDef {
    const char *name, *expr;
};
static const ConstDef kConstants[] = {
    {"PI", "(double)3.141592654"},
    {"ELECTRON_CHARGE", "-(double)1.602177e-19"},
    {"ELECTRON_CHARGE2", "(double)2.566971e-38"},
    {"CHARGE", "(double)1.602177e-19"},
    {"ELECTRON_MASS", "(double)9.10939e-31"},
    {"AMU", "(double)1.66054e-27"},
    {"K", "(double)1.38066e-23"},
    {"EPSILON", "(double)8.854e-12"},
    {"TRUE", "1"},
    {"FALSE", "0"},
    {"YES", "1"},
    {"NO", "0"},
};

struct Val {
    enum Kind { INT, DBL, STR } k = INT;
    long long i = 0;
    double d = 0.0;
    std::string s;
    double num() const { return k == INT ? (double)i : d; }
};

struct Evaluator {
    std::map<std::string, std::string> defs;
    std::map<std::string, Val> memo;
    std::map<std::string, int> busy;
    std::string err;
    uint64_t now_seed = 0;

    // --- tokenizer over one expression
    const char *p = nullptr;

    void ws() {
        while (*p && isspace((unsigned char)*p)) p++;
    }
    bool eat(char c) {
        ws();
        if (*p == c) {
            p++;
            return true;
        }
        return false;
    }
    bool fail(const std::string &m) {
        if (err.empty()) err = m;
        return false;
    }

    bool eval_macro(const std::string &name, Val &out) {
        auto m = memo.find(name);
        if (m != memo.end()) {
            out = m->second;
            return true;
        }
        auto d = defs.find(name);
        if (d == defs.end()) return fail("undefined macro " + name);
        if (busy[name]) return fail("recursive macro " + name);
        busy[name] = 1;
        const char *save = p;
        p = d->second.c_str();
        bool ok = expr(out);
        ws();
        if (ok && *p) ok = fail("trailing text in macro " + name + ": '" + std::string(p) + "'");
        p = save;
        busy[name] = 0;
        if (ok) memo[name] = out;
        return ok;
    }

    bool number(Val &v) {
        ws();
        const char *s = p;
        bool isdbl = false;
        while (isdigit((unsigned char)*p) || *p == '.' || *p == 'e' || *p == 'E' ||
               ((*p == '+' || *p == '-') && (p[-1] == 'e' || p[-1] == 'E'))) {
            if (*p == '.' || *p == 'e' || *p == 'E') isdbl = true;
            p++;
        }
        std::string t(s, p);
        while (*p == 'f' || *p == 'F' || *p == 'l' || *p == 'L' || *p == 'u' || *p == 'U') {
            if (*p == 'f' || *p == 'F') isdbl = true;
            p++;
        }
        if (isdbl) {
            v.k = Val::DBL;
            v.d = strtod(t.c_str(), nullptr);
        } else {
            v.k = Val::INT;
            v.i = strtoll(t.c_str(), nullptr, 10);
        }
        return true;
    }

    bool ident(std::string &id) {
        ws();
        if (!(isalpha((unsigned char)*p) || *p == '_')) return false;
        const char *s = p;
        while (isalnum((unsigned char)*p) || *p == '_') p++;
        id.assign(s, p);
        return true;
    }

    bool call(const std::string &f, Val &v) {
        // f '(' args ')'
        std::vector<Val> a;
        ws();
        if (!eat(')')) {
            for (;;) {
                Val x;
                if (f == "time") {  // time(NULL): the seed of a time-seeded run
                    std::string id;
                    if (!ident(id) && !number(x)) return fail("bad time() argument");
                    a.push_back(Val());
                } else {
                    if (!expr(x)) return false;
                    a.push_back(x);
                }
                if (eat(')')) break;
                if (!eat(',')) return fail("expected , or ) in call to " + f);
            }
        }
        v.k = Val::DBL;
        if (f == "time") {
            v.k = Val::INT;
            v.i = (long long)now_seed;
            return true;
        }
        if (a.empty()) return fail("missing argument to " + f);
        double x = a[0].num();
        if (f == "sqrt") v.d = sqrt(x);
        else if (f == "cos") v.d = cos(x);
        else if (f == "sin") v.d = sin(x);
        else if (f == "exp") v.d = exp(x);
        else if (f == "log") v.d = log(x);
        else if (f == "fabs" || f == "abs") v.d = fabs(x);
        else if (f == "pow" && a.size() == 2) v.d = pow(x, a[1].num());
        else return fail("unsupported function " + f);
        return true;
    }

    bool primary(Val &v) {
        ws();
        if (*p == '"') {
            p++;
            const char *s = p;
            while (*p && *p != '"') p++;
            if (!*p) return fail("unterminated string");
            v.k = Val::STR;
            v.s.assign(s, p);
            p++;
            return true;
        }
        if (isdigit((unsigned char)*p) || (*p == '.' && isdigit((unsigned char)p[1]))) return number(v);
        std::string id;
        if (ident(id)) {
            if (id == "true") {
                v.k = Val::INT;
                v.i = 1;
                return true;
            }
            if (id == "false" || id == "NULL") {
                v.k = Val::INT;
                v.i = 0;
                return true;
            }
            if (eat('(')) return call(id, v);
            return eval_macro(id, v);
        }
        if (eat('(')) {
            // cast or parenthesised expression
            const char *save = p;
            std::string t;
            if (ident(t)) {
                std::string t2 = t;
                ws();
                if (t == "unsigned" || t == "long") {
                    std::string more;
                    const char *s2 = p;
                    if (ident(more) && (more == "int" || more == "long")) t2 = "int";
                    else p = s2;
                    t2 = "int";
                }
                bool star = eat('*');
                if ((t == "double" || t == "float" || t == "int" || t == "char" || t == "unsigned" ||
                     t == "long" || t == "size_t") &&
                    eat(')')) {
                    Val x;
                    if (!unary(x)) return false;
                    if (t == "double") {
                        v.k = Val::DBL;
                        v.d = x.num();
                    } else if (t == "float") {
                        v.k = Val::DBL;
                        v.d = (double)(float)x.num();
                    } else if (t == "char" && star) {
                        v = x;
                    } else {
                        v.k = Val::INT;
                        v.i = (x.k == Val::INT) ? x.i : (long long)x.d;
                    }
                    return true;
                }
                (void)t2;
            }
            p = save;
            if (!expr(v)) return false;
            if (!eat(')')) return fail("expected )");
            return true;
        }
        return fail(std::string("unexpected text '") + p + "'");
    }

    bool unary(Val &v) {
        ws();
        if (*p == '-' || *p == '+') {
            char op = *p++;
            if (!unary(v)) return false;
            if (op == '-') {
                if (v.k == Val::INT) v.i = -v.i;
                else v.d = -v.d;
            }
            return true;
        }
        if (*p == '!') {
            p++;
            if (!unary(v)) return false;
            long long t = (v.num() == 0.0);
            v.k = Val::INT;
            v.i = t;
            return true;
        }
        return primary(v);
    }

    static void arith(Val &a, const Val &b, char op) {
        if (a.k == Val::INT && b.k == Val::INT) {
            switch (op) {
                case '*': a.i *= b.i; break;
                case '/': a.i = b.i ? a.i / b.i : 0; break;
                case '%': a.i = b.i ? a.i % b.i : 0; break;
                case '+': a.i += b.i; break;
                case '-': a.i -= b.i; break;
            }
            return;
        }
        double x = a.num(), y = b.num();
        a.k = Val::DBL;
        switch (op) {
            case '*': a.d = x * y; break;
            case '/': a.d = x / y; break;
            case '%': a.d = fmod(x, y); break;
            case '+': a.d = x + y; break;
            case '-': a.d = x - y; break;
        }
    }

    bool term(Val &v) {
        if (!unary(v)) return false;
        for (;;) {
            ws();
            char op = *p;
            if (op != '*' && op != '/' && op != '%') return true;
            p++;
            Val b;
            if (!unary(b)) return false;
            arith(v, b, op);
        }
    }

    bool expr(Val &v) {
        if (!term(v)) return false;
        for (;;) {
            ws();
            char op = *p;
            if (op != '+' && op != '-') return true;
            p++;
            Val b;
            if (!term(b)) return false;
            arith(v, b, op);
        }
    }
};

std::string strip_comments(const std::string &text) {
    std::string o;
    o.reserve(text.size());
    bool in_str = false;
    for (size_t k = 0; k < text.size(); k++) {
        char c = text[k];
        if (in_str) {
            o += c;
            if (c == '"') in_str = false;
            continue;
        }
        if (c == '"') {
            in_str = true;
            o += c;
            continue;
        }
        if (c == '/' && k + 1 < text.size() && text[k + 1] == '/') {
            while (k < text.size() && text[k] != '\n') k++;
            o += '\n';
            continue;
        }
        if (c == '/' && k + 1 < text.size() && text[k + 1] == '*') {
            k += 2;
            while (k + 1 < text.size() && !(text[k] == '*' && text[k + 1] == '/')) {
                if (text[k] == '\n') o += '\n';
                k++;
            }
            k++;
            o += ' ';
            continue;
        }
        o += c;
    }
    return o;
}

}  // namespace

// parameters.h:225-313 for the current node count; period = (int)PERIOD
static void pic_params_default_diagnostics(pic_params *P, int period) {
    const int N = P->nodes;
    P->graph_steps = 50 * period;
    P->graph_start = 250 * period;
    P->time_data_freq = 100;
    P->no_bins = 30001;
    P->divide_bins_by = (P->no_bins - 1) / 100;
    P->dist_loc[0] = (int)(0.9f * (float)N);
    P->dist_loc[1] = (int)(0.5f * (float)N);
    P->dist_loc[2] = (int)(0.1f * (float)N);
    P->dist_loc[3] = (int)(0.5f * (float)N);
    P->spread = (int)(0.05f * (float)N);
    P->probe[0] = 3;
    P->probe[1] = N / 2;
    P->probe[2] = N / 2;
    P->probe[3] = N / 2;
    P->probe[4] = N - 3;
    P->probe[5] = N / 2;
    P->write_graphs = 1;
}

// particle-count control as pic.cpp:257-327 sets it up from parameters.h:
// three velocity groups per species; on when the walls are not steady-state
static void pic_params_default_control(pic_params *P) {
    P->reduce = P->wall == PIC_WALL_STEADY ? 0 : 1;
    P->no_groups = 3;
    P->size_to_check[0] = 1000001;  // SIZE_TO_CHECK_E
    P->size_to_check[1] = 1000001;  // SIZE_TO_CHECK
    for (int s = 0; s < 2; s++)
        for (int g = 0; g < 8; g++) {
            P->group_limit[s][g] = 0;
            P->new_group_size[s][g] = 0;
            P->v_limit[s][g] = 0.0;
        }
    const double vte = P->vte, vti = P->vti;
    const double ve[3] = {2.0 * vte, 3.0 * vte, 1000.0 * vte};  // VE1, VE2, UPPER_V_LIMIT
    const double vi[3] = {2.0 * vti, 3.0 * vti, 1000.0 * vti};  // VI1, VI2, UPPER_V_LIMIT_I
    for (int g = 0; g < 3; g++) {
        P->v_limit[0][g] = ve[g];
        P->group_limit[0][g] = 400000;  // UPPER_GROUP_LIMIT
        P->new_group_size[0][g] = 200000;  // NEW_GROUP_SIZE
        P->v_limit[1][g] = vi[g];
        P->group_limit[1][g] = g == 0 ? 400000 : 100000;  // UPPER_GROUP_LIMIT(2)
        P->new_group_size[1][g] = g == 0 ? 200000 : 50000;  // NEW_GROUP_SIZE(2)
    }
}

extern "C" int pic_params_default(pic_params *P) {
    if (!P) return PIC_ERR_ARG;
    memset(P, 0, sizeof(*P));
    // constants.h
    const double PI = 3.141592654, CHARGE = 1.602177e-19, ELECTRON_CHARGE = -1.602177e-19,
                 ELECTRON_CHARGE2 = 2.566971e-38, ELECTRON_MASS = 9.10939e-31, AMU = 1.66054e-27,
                 EPSILON = 8.854e-12;
    // parameters.h defaults
    const double TEMPERATURE = 1.5, DENS = 1.0e15;
    P->nodes = 257;
    P->init_particle_no = 512 * 512;
    P->mode = PIC_MODE_REF;
    P->loader = PIC_LOAD_LATTICE;
    P->wall = PIC_WALL_STEADY;
    P->have_collisions = 0;
    P->gas = PIC_GAS_ARGON;
    P->sigma_cx = P->sigma_elastic = P->sigma_ionization = PIC_SIGMA_CONSTANT;
    P->mg_mode = PIC_MG_FIXED;
    P->mg_ncycle = 2;
    P->mg_max_iter = 30;
    P->mg_tol = 1e-8;
    P->sort_interval = 0;
    P->number_to_inject = 0;
    P->no_ion_species = 1;
    P->corner8 = 1;
    P->seed = (uint64_t)time(NULL);
    P->electron_mass = ELECTRON_MASS;
    P->ion_mass = 39.948 * AMU;
    P->electron_charge = ELECTRON_CHARGE;
    P->ion_charge = CHARGE;
    P->vte = sqrt(TEMPERATURE * CHARGE / ELECTRON_MASS);
    P->vti = sqrt(TEMPERATURE * CHARGE / P->ion_mass);
    P->vtn = sqrt(0.025 * CHARGE / P->ion_mass);
    P->sigma_load_e = P->vte / sqrt(2.0);  // particle.cpp:452-464 samples sigma = vt/sqrt(2)
    P->sigma_load_i = P->vti / sqrt(2.0);
    P->epsilon = EPSILON;
    double debye = sqrt((EPSILON * TEMPERATURE) / (DENS * CHARGE));
    P->length = (double)64 * debye;
    double wpe = sqrt((DENS * ELECTRON_CHARGE2) / (EPSILON * ELECTRON_MASS));
    P->dt = 0.1 / wpe;
    P->dti = 1.0 * P->dt;
    P->dt_dti = (int)(P->dti / P->dt);
    double two_pi_f_dt = 2.0 * P->dt * PI * 13.6e6;
    double period = 1.0 / two_pi_f_dt;
    P->run_time = (int)300 * (int)period + 5;
    P->qce = 1.0e6 * ELECTRON_CHARGE;
    P->qci = 1.0e6 * CHARGE;
    P->neutral_density = (0.2 * 133.32) / (0.025 * CHARGE);
    P->threshold = 15.75;
    P->magnetic_field = 0.0;
    P->angle = 30.0 * PI / 180.0;
    pic_params_default_diagnostics(P, (int)period);
    pic_params_default_control(P);
    // surface physics (parameters.h:91,141-152; pic.cpp:459-463)
    P->e_impact_see = P->ion_impact_see = P->reflection = P->backscatter = 0;
    P->see_max_yield = 1.4;
    P->see_E0 = 4.32;
    P->see_Emax0 = 650.0;
    P->work_function = 4.6;
    P->sort_key = -1;
    P->mg_sweeps = -1;
    P->null_bound = PIC_NULL_BOUND_BINNED;
    P->graphs = 0;
    return PIC_OK;
}

extern "C" int pic_params_load(const char *path, pic_params *P) {
    if (!path || !P) return PIC_ERR_ARG;
    std::ifstream in(path);
    if (!in) return pic_host::set_global_error(PIC_ERR_FILE, std::string("cannot open ") + path);
    std::stringstream ss;
    ss << in.rdbuf();
    std::string text = strip_comments(ss.str());
    Evaluator ev;
    ev.now_seed = (uint64_t)time(NULL);
    for (const ConstDef &c : kConstants) ev.defs[c.name] = c.expr;
    // join continuation lines, collect #define NAME body
    std::string line;
    std::istringstream ls(text);
    std::string acc;
    while (std::getline(ls, line)) {
        if (!line.empty() && line.back() == '\\') {
            acc += line.substr(0, line.size() - 1) + " ";
            continue;
        }
        acc += line;
        const char *s = acc.c_str();
        while (*s && isspace((unsigned char)*s)) s++;
        if (*s == '#') {
            s++;
            while (*s && isspace((unsigned char)*s)) s++;
            if (strncmp(s, "define", 6) == 0 && isspace((unsigned char)s[6])) {
                s += 6;
                while (*s && isspace((unsigned char)*s)) s++;
                const char *n0 = s;
                while (*s && (isalnum((unsigned char)*s) || *s == '_')) s++;
                std::string name(n0, s);
                if (*s == '(') {
                    acc.clear();
                    continue;  // function-like macros are not used by the configuration
                }
                std::string body(s);
                size_t a = body.find_first_not_of(" \t\r"), b = body.find_last_not_of(" \t\r");
                body = (a == std::string::npos) ? "" : body.substr(a, b - a + 1);
                if (!name.empty()) ev.defs[name] = body;
            }
        }
        acc.clear();
    }
    pic_params_default(P);
    P->seed = ev.now_seed;
    auto has = [&](const char *n) { return ev.defs.count(n) && !ev.defs[n].empty(); };
    auto num = [&](const char *n, double &dst) -> bool {
        if (!has(n)) return true;
        Val v;
        if (!ev.eval_macro(n, v) || v.k == Val::STR) return false;
        dst = v.num();
        return true;
    };
    auto inum = [&](const char *n, int32_t &dst) -> bool {
        double d;
        if (!has(n)) return true;
        if (!num(n, d)) return false;
        dst = (int32_t)d;
        return true;
    };
    auto str = [&](const char *n, std::string &dst) -> bool {
        if (!has(n)) return true;
        Val v;
        if (!ev.eval_macro(n, v)) return false;
        if (v.k != Val::STR) return false;
        dst = v.s;
        return true;
    };
    bool ok = true;
    ok = ok && inum("X_NO_OF_CELLS", P->nodes);
    int32_t ny = P->nodes;
    ok = ok && inum("Y_NO_OF_CELLS", ny);
    ok = ok && inum("INIT_PARTICLE_NO", P->init_particle_no);
    ok = ok && num("XLENGTH", P->length);
    double ylen = P->length;
    ok = ok && num("YLENGTH", ylen);
    ok = ok && num("DT", P->dt);
    ok = ok && num("DTI", P->dti);
    ok = ok && inum("DT_DTI", P->dt_dti);
    ok = ok && num("ELECTRON_MASS", P->electron_mass);
    ok = ok && num("ION_MASS", P->ion_mass);
    ok = ok && num("ELECTRON_CHARGE", P->electron_charge);
    ok = ok && num("ION_CHARGE", P->ion_charge);
    ok = ok && num("QCE", P->qce);
    ok = ok && num("QCI", P->qci);
    ok = ok && num("VTE", P->vte);
    ok = ok && num("VTI", P->vti);
    ok = ok && num("VTN", P->vtn);
    P->sigma_load_e = P->vte / sqrt(2.0);
    P->sigma_load_i = P->vti / sqrt(2.0);
    ok = ok && num("EPSILON", P->epsilon);
    ok = ok && num("NEUTRAL_DENSITY", P->neutral_density);
    ok = ok && num("THRESHOLD", P->threshold);
    ok = ok && num("SIGMA_VALUE", P->sigma_value);
    ok = ok && num("ELASTIC_SIGMA_VALUE", P->elastic_sigma_value);
    ok = ok && num("IONIZATION_SIGMA_VALUE", P->ionization_sigma_value);
    ok = ok && num("MAGNETIC_FIELD", P->magnetic_field);
    ok = ok && num("ANGLE", P->angle);
    ok = ok && num("PHI_LEFT", P->phi_left);
    ok = ok && num("PHI_RIGHT", P->phi_right);
    ok = ok && num("PHI_BOTTOM", P->phi_bottom);
    ok = ok && num("PHI_TOP", P->phi_top);
    ok = ok && num("PHI_PERIOD", P->phi_period);
    ok = ok && inum("TOTAL_RUN_TIME", P->run_time);
    ok = ok && inum("NUMBER_TO_INJECT", P->number_to_inject);
    ok = ok && inum("NO_ION_SPECIES", P->no_ion_species);
    double hc = P->have_collisions;
    ok = ok && num("HAVE_COLLISIONS", hc);
    P->have_collisions = hc != 0.0;
    double seed = 0;
    if (has("RAND_SEED")) {
        ok = ok && num("RAND_SEED", seed);
        P->seed = (uint64_t)(int64_t)seed;
    }
    std::string s;
    if (ok && has("SPECIES")) {
        ok = str("SPECIES", s);
        P->gas = (s == "HYDROGEN") ? PIC_GAS_HYDROGEN : PIC_GAS_ARGON;
    }
    auto sigma_mode = [&](const char *n, int32_t &dst) {
        if (!ok || !has(n)) return;
        std::string t;
        ok = str(n, t);
        dst = (t == "CONSTANT") ? PIC_SIGMA_CONSTANT : PIC_SIGMA_ENERGY;
    };
    sigma_mode("SIGMA", P->sigma_cx);
    sigma_mode("SIGMA_ELASTIC", P->sigma_elastic);
    sigma_mode("SIGMA_IONIZATION", P->sigma_ionization);
    if (ok && has("STEADYSTATE")) {
        ok = str("STEADYSTATE", s);
        P->wall = (s == "YES") ? PIC_WALL_STEADY : PIC_WALL_ABSORB;
    }
    // surface physics of the non-steady walls (parameters.h:136-152)
    auto yes = [&](const char *n, int32_t &dst) {
        if (!ok || !has(n)) return;
        std::string t;
        ok = str(n, t);
        dst = t == "YES";
    };
    yes("E_IMPACT_SEE", P->e_impact_see);
    yes("ION_IMPACT_SEE", P->ion_impact_see);
    yes("REFLECTION", P->reflection);
    yes("BACKSCATTER", P->backscatter);
    ok = ok && num("WORK_FUNCTION", P->work_function);
    // build extensions (absent from the reference): run-time options of this library
    if (ok && has("PIC_MODE")) {
        ok = str("PIC_MODE", s);
        P->mode = (s == "PHYSICAL") ? PIC_MODE_PHYSICAL : PIC_MODE_REF;
    }
    if (ok && has("PIC_LOADER")) {
        ok = str("PIC_LOADER", s);
        P->loader = (s == "RANDOM") ? PIC_LOAD_RANDOM : PIC_LOAD_LATTICE;
    }
    if (ok && has("PIC_MG")) {
        ok = str("PIC_MG", s);
        P->mg_mode = (s == "TOL") ? PIC_MG_TOL : PIC_MG_FIXED;
    }
    ok = ok && num("PIC_MG_TOL", P->mg_tol);
    ok = ok && inum("PIC_MG_NCYCLE", P->mg_ncycle);
    ok = ok && inum("PIC_MG_MAX_ITER", P->mg_max_iter);
    ok = ok && inum("PIC_MG_SWEEPS", P->mg_sweeps);
    ok = ok && inum("PIC_NULL_BOUND", P->null_bound);
    ok = ok && inum("PIC_GRAPHS", P->graphs);
    ok = ok && inum("PIC_SORT_INTERVAL", P->sort_interval);
    // diagnostic defaults follow the loaded grid and period; explicit macros win
    double period = 0.0;
    if (has("PERIOD")) ok = ok && num("PERIOD", period);
    else period = 1.0 / (2.0 * P->dt * 3.141592654 * 13.6e6);
    pic_params_default_diagnostics(P, (int)period);
    ok = ok && inum("START_GRAPH_TIME", P->graph_start);
    ok = ok && inum("NO_GRAPH_STEPS", P->graph_steps);
    ok = ok && inum("TIME_DATA_FREQUENCY", P->time_data_freq);
    ok = ok && inum("NO_BINS", P->no_bins);
    ok = ok && inum("DIVIDE_BINS_BY", P->divide_bins_by);
    ok = ok && inum("SPREAD", P->spread);
    ok = ok && inum("DIST_LOCATIONX", P->dist_loc[0]);
    ok = ok && inum("DIST_LOCATIONY", P->dist_loc[1]);
    ok = ok && inum("DIST_LOCATION2X", P->dist_loc[2]);
    ok = ok && inum("DIST_LOCATION2Y", P->dist_loc[3]);
    ok = ok && inum("POINT1X", P->probe[0]);
    ok = ok && inum("POINT1Y", P->probe[1]);
    ok = ok && inum("POINT2X", P->probe[2]);
    ok = ok && inum("POINT2Y", P->probe[3]);
    ok = ok && inum("POINT3X", P->probe[4]);
    ok = ok && inum("POINT3Y", P->probe[5]);
    // particle-count control (parameters.h:320-366) follows the loaded
    // thermal speeds and walls; explicit macros win
    pic_params_default_control(P);
    if (ok && has("NO_GROUPS")) {
        ok = inum("NO_GROUPS", P->no_groups);
        if (ok && P->no_groups != 3)
            return pic_host::set_global_error(
                PIC_ERR_UNSUPPORTED, "NO_GROUPS must be 3 (pic.cpp:257-327 sets three groups)");
    }
    ok = ok && inum("SIZE_TO_CHECK_E", P->size_to_check[0]);
    ok = ok && inum("SIZE_TO_CHECK", P->size_to_check[1]);
    int32_t ugl = P->group_limit[0][0], ugl2 = P->group_limit[1][1];
    int32_t ngs = P->new_group_size[0][0], ngs2 = P->new_group_size[1][1];
    ok = ok && inum("UPPER_GROUP_LIMIT", ugl);
    ok = ok && inum("UPPER_GROUP_LIMIT2", ugl2);
    ok = ok && inum("NEW_GROUP_SIZE", ngs);
    ok = ok && inum("NEW_GROUP_SIZE2", ngs2);
    ok = ok && num("VE1", P->v_limit[0][0]);
    ok = ok && num("VE2", P->v_limit[0][1]);
    ok = ok && num("UPPER_V_LIMIT", P->v_limit[0][2]);
    ok = ok && num("VI1", P->v_limit[1][0]);
    ok = ok && num("VI2", P->v_limit[1][1]);
    ok = ok && num("UPPER_V_LIMIT_I", P->v_limit[1][2]);
    for (int g = 0; g < 3; g++) {
        P->group_limit[0][g] = ugl;
        P->new_group_size[0][g] = ngs;
        P->group_limit[1][g] = g == 0 ? ugl : ugl2;
        P->new_group_size[1][g] = g == 0 ? ngs : ngs2;
    }
    if (ok && has("PIC_REDUCE")) ok = inum("PIC_REDUCE", P->reduce);
    ok = ok && num("PIC_SIGMA_LOAD_E", P->sigma_load_e);
    ok = ok && num("PIC_SIGMA_LOAD_I", P->sigma_load_i);
    if (!ok)
        return pic_host::set_global_error(PIC_ERR_FORMAT, "parameters file " + std::string(path) +
                                                              ": " + ev.err);
    if (ny != P->nodes || ylen != P->length)
        return pic_host::set_global_error(PIC_ERR_UNSUPPORTED,
                                          "X and Y grids must be identical (mglin needs a square grid)");
    return PIC_OK;
}
