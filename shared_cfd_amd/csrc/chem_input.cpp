// chem_input.cpp -- reads a Chemkin mechanism the way the reference does:
// read_cheminp (store/therm_lib/chemkin/core/sub_chem.f90:191-598) for
// chem.inp with the element weights of MW.inp, and set_trans_data
// (:788-889) for trans.inp.  Host code, no GPU.
//
// Conventions kept from the reference parser:
//  * lines are read 100 characters wide (next_line, :519-539): a '!' in
//    column 1 skips the line, a later '!' starts a comment, blank lines are
//    skipped;
//  * the section headers are "elements", "species", "thermo all" and
//    "reactions      cal/mole  moles"; sections end with "end"/"END";
//  * NASA-7 records are fixed-format: name in columns 1-24, four (a2,i3)
//    element counts, the phase in column 45 (must be G), Tlow/Thigh/Tcommon
//    in 2f10.0,f8.0 (zero Tlow: the defaults of the first line), then three
//    lines of 5E15 / 5E15 / 4E15 coefficient fields (high range first);
//  * a reaction line ends with A, beta, Ea; abe = (ln A, beta, Ea/Rc) with
//    Rc = 1.987 held in single precision (1d0/(1.987e0), mod_chem.raw.f90:9);
//    species on each side are separated by '+' with no blanks; "<=>" or "="
//    separate the sides, "=>" makes the reaction irreversible; "(+M)" marks a
//    fall-off reaction, "+M" a third-body one;
//  * auxiliary lines: REV /A b E/, LOW /A b E/, TROE /a T*** T* T**/ (stored
//    as a, 1/T***, 1/T*, T**), DUPLICATE, and "SP/eff/" efficiency lists
//    (stored as eff - 1);
//  * trans.inp: a title line, then per entry a header (A15,x,A15,3x,A1,I1,
//    A1,I1) -- entries with a second name (binary interactions) or unknown
//    species are skipped -- followed by its V lines (1x,a1,2(f7.1,2x),4e15.8)
//    and C lines (ignored).
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "../../include/shcfd.h"
#include "shcfd_nasa9.h"

namespace {

struct ParseError {
    std::string msg;
};

std::string rtrim(const std::string& s) {
    size_t e = s.find_last_not_of(' ');
    return e == std::string::npos ? std::string() : s.substr(0, e + 1);
}
std::string ltrim(const std::string& s) {
    size_t b = s.find_first_not_of(' ');
    return b == std::string::npos ? std::string() : s.substr(b);
}
std::string trim(const std::string& s) { return ltrim(rtrim(s)); }

// tabs count as blanks; lines are cut at 100 columns like the reference's a100 reads
std::string fixline(std::string s, size_t width) {
    if (!s.empty() && s.back() == '\r') s.pop_back();
    for (char& c : s)
        if (c == '\t') c = ' ';
    if (s.size() > width) s.resize(width);
    return s;
}

class Lines {
  public:
    explicit Lines(const char* path) : in_(path) {
        if (!in_) throw ParseError{std::string("cannot open ") + path};
    }
    // next_line (sub_chem.f90:519-539); "" at end of file
    std::string next() {
        std::string s;
        while (std::getline(in_, s)) {
            s = fixline(s, 100);
            size_t ind = s.find('!');
            if (ind == 0) continue;
            if (ind != std::string::npos) s = s.substr(0, ind);
            s = rtrim(s);
            if (!s.empty()) return s;
        }
        return std::string();
    }
    bool raw(std::string& s, size_t width) {
        if (!std::getline(in_, s)) return false;
        s = fixline(s, width);
        return true;
    }

  private:
    std::ifstream in_;
};

// Fortran list-directed real (D exponents allowed)
double to_real(std::string t, const char* what) {
    t = trim(t);
    for (char& c : t)
        if (c == 'd' || c == 'D') c = 'E';
    if (t.empty()) throw ParseError{std::string("missing number in ") + what};
    char* end = nullptr;
    double v = std::strtod(t.c_str(), &end);
    if (end == t.c_str() || *end != '\0') throw ParseError{std::string("bad number '") + t + "' in " + what};
    return v;
}

// fixed-width E/F field: blanks are ignored (BLANK='NULL'), an empty field is 0
double field(const std::string& line, size_t col0, size_t width) {
    std::string f = col0 < line.size() ? line.substr(col0, width) : std::string();
    std::string t;
    for (char c : f)
        if (c != ' ') t += (c == 'd' || c == 'D') ? 'E' : c;
    if (t.empty()) return 0.0;
    // Fortran allows an exponent without 'E' after a sign ("1.0+05")
    size_t p = t.find_first_of("+-", 1);
    if (p != std::string::npos && t[p - 1] != 'E' && t[p - 1] != 'e') t.insert(p, "E");
    char* end = nullptr;
    double v = std::strtod(t.c_str(), &end);
    if (end == t.c_str() || *end != '\0') throw ParseError{std::string("bad fixed-format field '") + f + "'"};
    return v;
}
int ifield(const std::string& line, size_t col0, size_t width) {
    std::string f = col0 < line.size() ? line.substr(col0, width) : std::string();
    f = trim(f);
    return f.empty() ? 0 : std::atoi(f.c_str());
}

// read_head_char (:560-573): first blank-delimited token
std::string head_token(std::string& line) {
    line = ltrim(line);
    size_t ind = line.find(' ');
    std::string t;
    if (ind != std::string::npos) {
        t = line.substr(0, ind);
        line = trim(line.substr(ind));
    } else {
        t = line;
        line.clear();
    }
    return t;
}
// read_tail_number (:550-559): last blank-delimited token as a real
double tail_number(std::string& line, const char* what) {
    line = rtrim(line);
    size_t ind = line.rfind(' ');
    std::string t = ind == std::string::npos ? line : line.substr(ind + 1);
    double v = to_real(t, what);
    line = ind == std::string::npos ? std::string() : rtrim(line.substr(0, ind + 1));
    return v;
}
// read_tail_number_arr (:504-518) with its unit conversions
const double INV_RC = 1.0 / (double)1.987f;
void tail_numbers(std::string& line, int num, double* arr, const char* what) {
    for (int i = num - 1; i >= 0; i--) arr[i] = tail_number(line, what);
    if (num == 3) {
        arr[0] = std::log(arr[0]);
        arr[2] = arr[2] * INV_RC;
    } else if (num == 4) {
        arr[1] = 1.0 / arr[1];
        arr[2] = 1.0 / arr[2];
    }
}
// fetch_slash (:588-597): text between the first two '/', rest after them
std::string fetch_slash(std::string& line) {
    size_t a = line.find('/');
    size_t b = a == std::string::npos ? std::string::npos : line.find('/', a + 1);
    if (b == std::string::npos) throw ParseError{"unbalanced '/' in '" + line + "'"};
    std::string s = line.substr(a + 1, b - a - 1);
    line = trim(line.substr(b + 1));
    return s;
}
// search_SYM (:574-587): pops the species before the first '+'
int pop_species(std::string& str, const std::vector<std::string>& sym) {
    size_t ind = str.find('+');
    std::string buf;
    if (ind == std::string::npos) {
        buf = rtrim(str);
        str.clear();
    } else {
        buf = rtrim(str.substr(0, ind));
        str = rtrim(str.substr(ind + 1));
    }
    for (size_t i = 0; i < sym.size(); i++)
        if (sym[i] == buf) return (int)i;
    throw ParseError{"unknown species '" + buf + "'"};
}

void parse(const char* chem_inp, const char* mw_inp, const char* trans_inp, shcfd_mech& m) {
    std::memset(&m, 0, sizeof(m));
    Lines L(chem_inp);
    std::vector<std::string> elm, spc;
    std::string line = L.next();
    if (line != "elements") throw ParseError{"elements declaration not found."};
    for (;;) {
        line = trim(L.next());
        if (line == "end" || line == "END") break;
        if (line.empty()) throw ParseError{"unexpected end of file in elements"};
        while (!line.empty()) {
            elm.push_back(head_token(line));
            if ((int)elm.size() > SHCFD_CHEM_MAX_NE) throw ParseError{"chem.inp element number exceeded assigned one."};
        }
    }
    m.ne = (int)elm.size();
    // element weights (MW.inp)
    std::vector<double> mwe(elm.size(), 0.0);
    std::vector<int> found(elm.size(), 0);
    {
        Lines M(mw_inp);
        for (;;) {
            std::string l = M.next();
            if (l.empty()) break;
            std::string e = head_token(l);
            for (size_t i = 0; i < elm.size(); i++)
                if (e == elm[i] && !found[i]) {
                    found[i] = 1;
                    mwe[i] = tail_number(l, "MW.inp");
                    break;
                }
        }
        for (int f : found)
            if (!f) throw ParseError{"Error. There is Element which does not exit in MW.inp."};
    }
    line = L.next();
    if (line != "species") throw ParseError{"species declaration not found."};
    for (;;) {
        line = trim(L.next());
        if (line == "end" || line == "END") break;
        if (line.empty()) throw ParseError{"unexpected end of file in species"};
        while (!line.empty()) {
            spc.push_back(head_token(line));
            if ((int)spc.size() > SHCFD_CHEM_MAX_NS) throw ParseError{"chem.inp species number exceeded assigned one."};
        }
    }
    m.ns = (int)spc.size();
    line = L.next();
    if (line != "thermo all") throw ParseError{"'thermo all' declaration not found."};
    line = L.next();
    double tdef[3] = {field(line, 0, 10), field(line, 10, 10), field(line, 20, 10)};
    std::vector<int> have(spc.size(), 0);
    for (;;) {
        line = L.next();
        if (line == "end" || line == "END") break;
        if (line.empty()) throw ParseError{"unexpected end of file in thermo"};
        std::string name = line.substr(0, std::min<size_t>(24, line.size()));
        size_t sp = name.find(' ');
        if (sp != std::string::npos) name = name.substr(0, sp);
        int k = -1;
        for (size_t i = 0; i < spc.size(); i++)
            if (spc[i] == name) k = (int)i;
        if (k < 0) {
            L.next();
            L.next();
            L.next();
            continue;
        }
        if (line.size() < 45 || line[44] != 'G') throw ParseError{"Error! Other than Gas phase is contained."};
        have[k] = 1;
        auto add_elem = [&](const std::string& sym2, int cnt) {
            std::string s = trim(sym2);
            if (s.empty()) return;
            for (size_t e = 0; e < elm.size(); e++)
                if (elm[e] == s) {
                    m.es[k][e] = cnt;
                    return;
                }
            throw ParseError{"unknown element '" + s + "' in species " + name};
        };
        for (int i = 0; i < 4; i++)
            add_elem(line.size() > (size_t)(24 + 5 * i) ? line.substr(24 + 5 * i, 2) : std::string(),
                     ifield(line, 26 + 5 * i, 3));
        double tb[3] = {field(line, 45, 10), field(line, 55, 10), field(line, 65, 8)};
        add_elem(line.size() > 73 ? line.substr(73, 2) : std::string(), ifield(line, 75, 3));
        double mw = 0.0;
        for (size_t e = 0; e < elm.size(); e++) mw = mw + m.es[k][e] * mwe[e];
        m.mw[k] = mw;
        for (int i = 0; i < 3; i++) m.tthre[k][i] = (tb[0] != 0.0) ? tb[i] : tdef[i];
        std::string l2 = L.next(), l3 = L.next(), l4 = L.next();
        for (int i = 0; i < 5; i++) m.coeff[k][1][i] = field(l2, 15 * i, 15);
        m.coeff[k][1][5] = field(l3, 0, 15);
        m.coeff[k][1][6] = field(l3, 15, 15);
        for (int i = 0; i < 3; i++) m.coeff[k][0][i] = field(l3, 30 + 15 * i, 15);
        for (int i = 0; i < 4; i++) m.coeff[k][0][3 + i] = field(l4, 15 * i, 15);
    }
    for (int h : have)
        if (!h) throw ParseError{"There is species which does not have thermal data."};
    line = L.next();
    if (line != "reactions      cal/mole  moles") throw ParseError{"'reactions' declaration not found."};
    int nr = 0;
    for (;;) {
        line = trim(L.next());
        if (line == "end" || line == "END") break;
        if (line.empty()) throw ParseError{"unexpected end of file in reactions"};
        if (line == "duplicate" || line == "DUPLICATE") continue;
        size_t slash = line.find('/');
        if (slash == std::string::npos) {
            if (nr >= SHCFD_CHEM_MAX_NR) throw ParseError{"chem.inp reaction number exceeded assigned one."};
            int r = nr++;
            tail_numbers(line, 3, m.abe[r], "reaction line");
            std::string lr, lp;
            size_t ind = line.find("<=>");
            if (ind != std::string::npos) {
                lr = line.substr(0, ind);
                lp = line.substr(ind + 3);
            } else {
                ind = line.find('=');
                if (ind == std::string::npos) throw ParseError{"no '=' in reaction '" + line + "'"};
                lr = line.substr(0, ind);
                lp = line.substr(ind + 1);
            }
            lr = rtrim(lr);
            lp = rtrim(lp);
            if (!lp.empty() && lp[0] == '>') {
                m.reverse_off[r] = 1;
                lp = lp.substr(1);
            }
            size_t im = lr.find("(+M)");
            if (im != std::string::npos) {
                if (im + 4 != lr.size()) throw ParseError{"Error. (+M) must be end of each side."};
                lr = lr.substr(0, lr.size() - 4);
                if (lp.size() < 4) throw ParseError{"Error. (+M) must be end of each side."};
                lp = lp.substr(0, lp.size() - 4);
                m.exist_m[r] = 1;
            } else {
                im = lr.find("+M");
                if (im != std::string::npos) {
                    if (im + 2 != lr.size()) throw ParseError{"Error. +M must be end of each side."};
                    lr = lr.substr(0, lr.size() - 2);
                    if (lp.size() < 2) throw ParseError{"Error. +M must be end of each side."};
                    lp = lp.substr(0, lp.size() - 2);
                    m.exist_m[r] = 1;
                    m.kind[r] = SHCFD_RXN_M;
                }
            }
            for (int side = 0; side < 2; side++) {
                std::string s = side == 0 ? lr : lp;
                int i = 0;
                for (;;) {
                    if (i >= 3) throw ParseError{"more than 3 species on one side of a reaction"};
                    m.ind_nu[r][side][i] = pop_species(s, spc);
                    i++;
                    if (s.empty()) break;
                }
                m.num_nu[r][side] = i;
            }
        } else {
            if (nr == 0) throw ParseError{"auxiliary line before the first reaction"};
            int r = nr - 1;
            std::string key = rtrim(line.substr(0, slash));
            std::string rest = line.substr(slash);
            if (key == "REV" || key == "rev") {
                if (m.reverse_off[r] != 0 || (m.kind[r] != SHCFD_RXN_NONE && m.kind[r] != SHCFD_RXN_M))
                    throw ParseError{"Error at rev"};
                m.kind[r] = m.kind[r] + 1;
                std::string s = fetch_slash(rest);
                tail_numbers(s, 3, m.cabe[r], "REV");
            } else if (key == "LOW" || key == "low") {
                if (m.reverse_off[r] != 0 || m.kind[r] != SHCFD_RXN_NONE) throw ParseError{"Error at low"};
                m.kind[r] = SHCFD_RXN_LOW;
                std::string s = fetch_slash(rest);
                tail_numbers(s, 3, m.cabe[r], "LOW");
            } else if (key == "TROE" || key == "troe") {
                if (m.reverse_off[r] != 0 || m.kind[r] != SHCFD_RXN_LOW) throw ParseError{"Error at TROE"};
                m.kind[r] = SHCFD_RXN_TROE;
                std::string s = fetch_slash(rest);
                tail_numbers(s, 4, m.troe[r], "TROE");
            } else {
                if (!m.exist_m[r]) throw ParseError{"Error at M enhanced"};
                for (int k = 0; k < SHCFD_CHEM_MAX_NS; k++) m.men[r][k] = 0.0;
                int j = 0;
                std::string sp = key;
                for (;;) {
                    if (j >= SHCFD_CHEM_MAX_NS) throw ParseError{"too many efficiencies"};
                    std::string tmp = sp;
                    m.ind_m[r][j] = pop_species(tmp, spc);
                    std::string s = fetch_slash(rest);
                    m.men[r][j] = tail_number(s, "efficiency") - 1.0;
                    j++;
                    if (rest.empty()) break;
                    size_t ind = rest.find('/');
                    if (ind == std::string::npos) throw ParseError{"bad efficiency list"};
                    sp = rest.substr(0, ind);
                    rest = rest.substr(ind);
                }
                m.num_m[r] = j;
            }
        }
    }
    for (int i = 0; i < nr; i++) m.snu[i] = m.num_nu[i][1] - m.num_nu[i][0];
    for (int j = 0; j < m.ns; j++) m.inv_mw[j] = 1.0 / m.mw[j];
    m.nr = nr > 1 ? nr : 1;

    // transport (set_trans_data)
    m.nt = 0;
    if (!trans_inp) return;
    std::ifstream T(trans_inp);
    if (!T) throw ParseError{std::string("cannot open ") + trans_inp};
    std::string s;
    if (!std::getline(T, s)) return;
    while (std::getline(T, s)) {
        s = fixline(s, 300);
        if (!s.empty() && s[0] == '!') continue;
        if (s.compare(0, 3, "END") == 0 || s.compare(0, 3, "end") == 0) break;
        std::string sname = rtrim(s.substr(0, std::min<size_t>(15, s.size())));
        std::string comment = s.size() > 16 ? trim(s.substr(16, 15)) : std::string();
        int iv = ifield(s, 35, 1), ic = ifield(s, 37, 1);
        int k = -1;
        for (size_t i = 0; i < spc.size(); i++)
            if (spc[i] == sname) k = (int)i;
        if (!comment.empty() || k < 0) {
            for (int i = 0; i < iv + ic; i++) std::getline(T, s);
            continue;
        }
        if (m.nt >= SHCFD_CHEM_MAX_NS || iv > SHCFD_CHEM_MAX_TSEC) throw ParseError{"trans.inp: too many entries"};
        int t = m.nt++;
        m.tr2th[t] = k;
        m.nsec_trans[t] = iv;
        for (int i = 0; i < iv; i++) {
            if (!std::getline(T, s)) throw ParseError{"ERROR : ODD FORMAT AT trans.inp"};
            s = fixline(s, 300);
            if (s.size() < 2 || s[1] != 'V') throw ParseError{"ERROR : ODD FORMAT AT trans.inp"};
            m.trange[t][i][0] = field(s, 2, 7);
            m.trange[t][i][1] = field(s, 11, 7);
            for (int j = 0; j < 4; j++) m.trans[t][i][j] = field(s, 20 + 15 * j, 15);
        }
        for (int i = 0; i < ic; i++) {
            if (!std::getline(T, s)) throw ParseError{"ERROR : ODD FORMAT AT trans.inp"};
            if (s.size() < 2 || s[1] != 'C') throw ParseError{"ERROR : ODD FORMAT AT trans.inp"};
        }
    }
}

// ---------------------------------------------------------------------------
// NASA-9 tables of the cold-flow model: read_cheminp (NASA/core/
// sub_chem.f90:2-67: the element and species lists of chem.inp),
// set_therm_data (:68-166: CEA thermo.inp -- comment lines '!', a title line,
// the default-ranges line, then per species a name line (A18), a line
// (i2,8x,5(a2,f6.2),2x,f13.7) with the section count and the molecular
// weight, and per section the range (2(1x,f10.3)) and the coefficients
// (5D16.9 / 2D16.8,16x,2D16.8); reading stops at the first "END"),
// set_trans_data (:167-245, as for the Chemkin model) and
// read_fo_composition (:247-300: pressure, temperature and mole ratios of
// the oxidizer, then of the fuel, from control_chem.inp), normalised to
// mol/kg as initialize_cold_flow does (calc_ini, :420-441).
// ---------------------------------------------------------------------------
void parse_nasa9(const char* chem_inp, const char* thermo_inp, const char* trans_inp, const char* control_inp,
                 shcfd_nasa9& m, double* pt) {
    std::memset(&m, 0, sizeof(m));
    std::vector<std::string> spc;
    {
        Lines L(chem_inp);
        std::string line = L.next();  // "elements"
        std::string buf;
        auto word = [&]() {
            if (buf.empty()) buf = L.next();
            if (buf.empty()) throw ParseError{"chem.inp: unexpected end of file"};
            return head_token(buf);
        };
        for (;;) {
            std::string w = word();
            if (w == "end" || w == "END") break;
        }
        line = buf.empty() ? L.next() : buf;  // "species"
        buf.clear();
        for (;;) {
            std::string w = word();
            if (w == "end" || w == "END") break;
            spc.push_back(w);
        }
    }
    if (spc.empty() || spc.size() > SHCFD_NASA9_MAX_NS) throw ParseError{"chem.inp: 1 to 8 species expected"};
    m.ns = (int)spc.size();
    auto find = [&](const std::string& n) {
        for (size_t i = 0; i < spc.size(); i++)
            if (spc[i] == n) return (int)i;
        return -1;
    };
    {
        std::ifstream T(thermo_inp);
        if (!T) throw ParseError{std::string("cannot open ") + thermo_inp};
        std::string s;
        do {
            if (!std::getline(T, s)) throw ParseError{"thermo.inp: no data"};
        } while (!s.empty() && s[0] == '!');
        std::getline(T, s);  // default ranges
        int found = 0;
        std::vector<int> seen(spc.size(), 0);
        while (found < m.ns && std::getline(T, s)) {
            s = fixline(s, 200);
            std::string name = trim(s.substr(0, std::min<size_t>(18, s.size())));
            if (s.compare(0, 3, "END") == 0) break;
            std::string l2;
            if (!std::getline(T, l2)) throw ParseError{"thermo.inp: truncated entry " + name};
            l2 = fixline(l2, 200);
            const int nsec = ifield(l2, 0, 2);
            const int j = find(name);
            if (j < 0 || seen[j]) {
                for (int k = 0; k < 3 * nsec; k++) std::getline(T, s);
                continue;
            }
            if (nsec < 1 || nsec > SHCFD_NASA9_MAX_SEC) throw ParseError{"thermo.inp: bad section count for " + name};
            seen[j] = 1;
            found++;
            m.nsec[j] = nsec;
            m.mw[j] = field(l2, 52, 13);
            for (int l = 0; l < nsec; l++) {
                std::string a, b, c;
                if (!std::getline(T, a) || !std::getline(T, b) || !std::getline(T, c))
                    throw ParseError{"thermo.inp: truncated entry " + name};
                a = fixline(a, 200);
                b = fixline(b, 200);
                c = fixline(c, 200);
                m.trange[j][l][0] = field(a, 1, 10);
                m.trange[j][l][1] = field(a, 12, 10);
                for (int k = 0; k < 5; k++) m.co[j][l][k] = field(b, 16 * k, 16);
                m.co[j][l][5] = field(c, 0, 16);
                m.co[j][l][6] = field(c, 16, 16);
                m.co[j][l][7] = field(c, 48, 16);
                m.co[j][l][8] = field(c, 64, 16);
            }
        }
        if (found != m.ns) throw ParseError{"Some species have not found in thermo.inp."};
    }
    if (trans_inp) {
        std::ifstream T(trans_inp);
        if (!T) throw ParseError{std::string("cannot open ") + trans_inp};
        std::string s;
        std::getline(T, s);
        while (std::getline(T, s)) {
            s = fixline(s, 300);
            if (!s.empty() && s[0] == '!') continue;
            if (s.compare(0, 3, "END") == 0 || s.compare(0, 3, "end") == 0) break;
            std::string sname = rtrim(s.substr(0, std::min<size_t>(15, s.size())));
            std::string comment = s.size() > 16 ? trim(s.substr(16, 15)) : std::string();
            int iv = ifield(s, 35, 1), ic = ifield(s, 37, 1);
            int k = find(trim(sname));
            if (!comment.empty() || k < 0) {
                for (int i = 0; i < iv + ic; i++) std::getline(T, s);
                continue;
            }
            if (m.nt >= SHCFD_NASA9_MAX_NS || iv > SHCFD_NASA9_MAX_TSEC || iv < 1)
                throw ParseError{"trans.inp: too many entries"};
            int t = m.nt++;
            m.tr2th[t] = k;
            m.nsec_trans[t] = iv;
            for (int i = 0; i < iv; i++) {
                if (!std::getline(T, s)) throw ParseError{"ERROR : ODD FORMAT AT trans.inp"};
                s = fixline(s, 300);
                if (s.size() < 2 || s[1] != 'V') throw ParseError{"ERROR : ODD FORMAT AT trans.inp"};
                m.trange_trans[t][i][0] = field(s, 2, 7);
                m.trange_trans[t][i][1] = field(s, 11, 7);
                for (int j = 0; j < 4; j++) m.trans[t][i][j] = field(s, 20 + 15 * j, 15);
            }
            for (int i = 0; i < ic; i++) {
                if (!std::getline(T, s)) throw ParseError{"ERROR : ODD FORMAT AT trans.inp"};
                if (s.size() < 2 || s[1] != 'C') throw ParseError{"ERROR : ODD FORMAT AT trans.inp"};
            }
        }
    }
    {
        std::ifstream C(control_inp);
        if (!C) throw ParseError{std::string("cannot open ") + control_inp};
        std::string s;
        double* comp[2] = {m.no, m.nf};
        double p_t[4] = {0.0, 0.0, 0.0, 0.0};
        std::getline(C, s);  // title
        for (int st = 0; st < 2; st++) {
            for (int k = 0; k < 2; k++) {
                if (!std::getline(C, s)) throw ParseError{"control_chem.inp: truncated"};
                s = fixline(s, 200);
                p_t[2 * st + k] = field(s, 25, 15);
            }
            std::getline(C, s);  // "... Composition(mole ratio):"
            for (;;) {
                if (!std::getline(C, s)) throw ParseError{"control_chem.inp: truncated composition"};
                s = trim(fixline(s, 200));
                if (s.compare(0, 3, "end") == 0) break;
                std::string rest = s;
                std::string name = head_token(rest);
                if (rest.empty()) throw ParseError{"Bad format at control_chem.inp while reading Compositions"};
                const int j = find(name);
                if (j < 0) throw ParseError{"control_chem.inp: unknown species " + name};
                comp[st][j] = to_real(rest, "control_chem.inp");
            }
        }
        for (int st = 0; st < 2; st++) {  // mole ratios -> mol/kg (calc_ini)
            double sm = 0.0;
            for (int j = 0; j < m.ns; j++) sm = sm + (comp[st][j] * m.mw[j]);
            for (int j = 0; j < m.ns; j++) comp[st][j] = (comp[st][j] / sm) * 1e3;
        }
        if (pt) {  // (po, To, pf, Tf)
            for (int k = 0; k < 4; k++) pt[k] = p_t[k];
        }
    }
}

}  // namespace

extern "C" int shcfd_nasa9_parse(const char* chem_inp, const char* thermo_inp, const char* trans_inp,
                                 const char* control_chem_inp, shcfd_nasa9* out, double* pt, char* errbuf,
                                 int errlen) {
    if (!chem_inp || !thermo_inp || !control_chem_inp || !out) return SHCFD_E_ARG;
    try {
        parse_nasa9(chem_inp, thermo_inp, trans_inp, control_chem_inp, *out, pt);
    } catch (const ParseError& e) {
        if (errbuf && errlen > 0) std::snprintf(errbuf, (size_t)errlen, "%s", e.msg.c_str());
        return SHCFD_E_ARG;
    }
    return SHCFD_OK;
}

// initialize_cold_flow for one stream (NASA/core/sub_chem.f90:388-488):
// E = Ru T sum n_j (h_j/(Ru T) - 1) at (p, T) (calc_ini, from E = 0), then
// cold_flow from T, rho = p / (Ru 1e3 / MW T) and set_static_qw: q = (Y rho,
// 0, 0, rho E), w = (rho, 0, 0, p, Y, kappa, E + p/rho, p/(rho T), mu), vhi.
extern "C" int shcfd_nasa9_stream(const shcfd_nasa9* m, int which, double p, double T, double* q, double* w,
                                  double* vhi) {
    if (!m || (which != 0 && which != 1) || !(p > 0.0) || !(T > 0.0) || !q || !w || !vhi) return SHCFD_E_ARG;
    if (m->ns < 1 || m->ns > SHCFD_NASA9_MAX_NS) return SHCFD_E_ARG;
    using namespace shcfd::nasa9;
    const double* n = which == 0 ? m->nf : m->no;
    double vh[9];
    v_thrt(T, shcfd::chem::dlog(T), vh);
    double E = 0.0;
    for (int j = 0; j < m->ns; j++) {
        const double* co = m->co[j][section(m, j, T)];
        double tmp = -1.0;
        for (int k = 0; k < 9; k++) tmp = tmp + (vh[k] * co[k]);
        E = E + (tmp * n[j]);
    }
    E = (E * RU) * T;
    const double Y[2] = {which == 0 ? 1.0 : 0.0, which == 0 ? 0.0 : 1.0};
    double Tn = T;
    ColdOut o;
    cold_flow(m, Y[0], Y[1], E, Tn, o);
    const double rho = p / (((RU * 1e3) / o.mwave) * Tn);
    for (int i = 0; i < 2; i++) {
        q[i] = Y[i] * rho;
        w[4 + i] = Y[i];
        vhi[i] = o.vhi[i];
    }
    q[2] = 0.0;
    q[3] = 0.0;
    q[4] = rho * E;
    w[0] = rho;
    w[1] = 0.0;
    w[2] = 0.0;
    w[3] = p;
    w[6] = o.kappa;
    w[7] = E + (p / rho);
    w[8] = p / (rho * Tn);
    w[9] = o.mu;
    return SHCFD_OK;
}

extern "C" int shcfd_mech_parse(const char* chem_inp, const char* mw_inp, const char* trans_inp, shcfd_mech* out,
                                char* errbuf, int errlen) {
    if (!chem_inp || !mw_inp || !out) return SHCFD_E_ARG;
    try {
        parse(chem_inp, mw_inp, trans_inp, *out);
    } catch (const ParseError& e) {
        if (errbuf && errlen > 0) std::snprintf(errbuf, (size_t)errlen, "%s", e.msg.c_str());
        return SHCFD_E_ARG;
    }
    return SHCFD_OK;
}
