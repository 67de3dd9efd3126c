// Molecular structures on the host: PDB reading, DFIRE atom typing, the
// normal-mode mask, centring, the minimum-volume-ellipsoid reference point
// and PDB writing.  These are the set-up and output steps around the GSO
// path (source/prep/src/read_input_structure.cc, complex.cc,
// initialization.cc, translate.cc; source/dock/src/prepare_gso_tasks.cc
// get_docking_model and minimum_volume_ellipsoid, write_pdb).
#include <algorithm>
#include <cctype>
#include <stdexcept>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "sd_internal.h"

namespace sd {

struct AtomRec {
    std::string record;  // columns 1-6 as read ("ATOM  ", "HETATM")
    int serial = 0;
    std::string name, altloc, resname, chain, insertion, element;
    int resseq = 0;
    double x = 0, y = 0, z = 0;  // PDB values rounded through float, as the reference stores them
    double occupancy = 1, bfactor = 0;
};

struct Residue {
    std::string name;
    int number = 0;
    std::string insertion;
    int first = 0, count = 0;  // atoms [first, first + count)
};

struct Structure {
    std::vector<AtomRec> atoms;
    std::vector<Residue> residues;
    std::vector<std::string> chains;
};

}  // namespace sd

struct sdgso_structure_s : sd::Structure {};

namespace sd {

static std::string trim(const std::string &s) {
    size_t a = 0, b = s.size();
    while (a < b && std::isspace((unsigned char)s[a])) a++;
    while (b > a && std::isspace((unsigned char)s[b - 1])) b--;
    return s.substr(a, b - a);
}

static std::string field(const std::string &line, size_t pos, size_t len) {
    if (pos >= line.size()) return std::string();
    return line.substr(pos, len);
}

// Elements recognised by the reference (atom.h:10), two-letter names first.
static bool known_element(const std::string &e) {
    static const char *kElems[] = {"C", "N", "O", "H", "S", "P", "CL", "MG", "FE", "PB", "SE", "F"};
    for (const char *k : kElems)
        if (e == k) return true;
    return false;
}

// A number field as the reference reads it: stof over the whole field, the
// default when nothing parses or characters (trailing blanks too) remain.
static double number_or(const std::string &s, double dflt) {
    char *end = nullptr;
    const float v = std::strtof(s.c_str(), &end);
    if (end == s.c_str() || (size_t)(end - s.c_str()) < s.size()) return dflt;
    return (double)v;
}

// One ATOM/HETATM line (read_input_structure.cc:131-218).  Returns false for
// an ignored atom; throws on an element the reference does not recognise.
static bool parse_atom(const std::string &line, bool no_oxt, bool no_h, bool no_water, AtomRec &a) {
    if (line.size() < 54) throw std::runtime_error("short ATOM/HETATM line: " + line);
    a.x = (double)(float)std::stod(line.substr(30, 8));
    a.y = (double)(float)std::stod(line.substr(38, 8));
    a.z = (double)(float)std::stod(line.substr(46, 8));
    a.serial = std::stoi(line.substr(6, 5));
    a.name = trim(line.substr(12, 4));
    a.altloc = trim(line.substr(16, 1));
    a.resname = trim(line.substr(17, 4));
    a.chain = trim(line.substr(21, 1));
    a.insertion = field(line, 26, 1);
    a.resseq = std::stoi(line.substr(22, 4));
    a.record = line.substr(0, 6);
    bool valid = true;
    if (no_water && a.resname == "HOH") valid = false;
    if ((no_h && !a.name.empty() && a.name[0] == 'H') || (no_oxt && a.name == "OXT")) valid = false;
    a.occupancy = number_or(field(line, 54, 6), 1.0);
    a.bfactor = number_or(field(line, 60, 6), 0.0);
    std::string e = a.name.substr(0, 2);
    if (!known_element(e)) {
        e = a.name.substr(0, 1);
        if (!known_element(e))
            throw std::runtime_error("atom " + std::to_string(a.serial) + " (" + a.name +
                                     "): element not recognised");
    }
    a.element = trim(e);
    return valid;
}

// The file's first model (read_input_structure.cc:221-307): residues start
// where (name, number, insertion) changes, chains where the chain id changes.
static void read_pdb(const std::string &path, bool no_oxt, bool no_h, bool no_water, Structure &s) {
    std::ifstream in(path);
    if (!in.is_open()) throw std::runtime_error("cannot open " + path);
    std::string line, last_chain = "#", last_res = "#", last_ins = "#";
    int last_num = 0, models = 0;
    bool have_res = false;
    while (std::getline(in, line)) {
        if (!line.empty() && line.back() == '\r') line.pop_back();
        if (models > 1) break;
        const std::string type = trim(field(line, 0, 6));
        if (type == "MODEL") {
            models++;
            continue;
        }
        if (type != "ATOM" && type != "HETATM") continue;
        AtomRec a;
        if (!parse_atom(line, no_oxt, no_h, no_water, a)) continue;
        if (a.chain != last_chain) {
            last_chain = a.chain;
            s.chains.push_back(a.chain);
        }
        if (!have_res || a.resname != last_res || a.resseq != last_num || a.insertion != last_ins) {
            last_res = a.resname;
            last_num = a.resseq;
            last_ins = a.insertion;
            have_res = true;
            Residue r;
            r.name = a.resname;
            r.number = a.resseq;
            r.insertion = a.insertion;
            r.first = (int)s.atoms.size();
            s.residues.push_back(r);
        }
        s.residues.back().count++;
        s.atoms.push_back(a);
    }
    if (s.atoms.empty()) throw std::runtime_error("no atoms read from " + path);
}

// DFIRE atom types (fastdfire.h: RES_3, atomnumber, atom_res_trans; applied
// by get_docking_model, prepare_gso_tasks.cc:31-69).  Every residue's atoms
// take consecutive type numbers in the listed order, from `first`.
struct DfireResidue {
    const char *name;
    int first;
    const char *atoms;  // space-separated, in type order
};
static const DfireResidue kDfire[] = {
    {"ALA", 74, "N CA C O CB"},
    {"CYS", 0, "N CA C O CB SG"},
    {"ASP", 122, "N CA C O CB CG OD1 OD2"},
    {"GLU", 113, "N CA C O CB CG CD OE1 OE2"},
    {"PHE", 14, "N CA C O CB CG CD1 CD2 CE1 CE2 CZ"},
    {"GLY", 79, "N CA C O"},
    {"HIS", 130, "N CA C O CB CG ND1 CD2 CE1 NE2"},
    {"ILE", 25, "N CA C O CB CG1 CG2 CD1"},
    {"LYS", 151, "N CA C O CB CG CD CE NZ"},
    {"LEU", 33, "N CA C O CB CG CD1 CD2"},
    {"MET", 6, "N CA C O CB CG SD CE"},
    {"ASN", 105, "N CA C O CB CG OD1 ND2"},
    {"PRO", 160, "N CA C O CB CG CD"},
    {"GLN", 96, "N CA C O CB CG CD OE1 NE2"},
    {"ARG", 140, "N CA C O CB CG CD NE CZ NH1 NH2"},
    {"SER", 90, "N CA C O CB OG"},
    {"THR", 83, "N CA C O CB OG1 CG2"},
    {"VAL", 41, "N CA C O CB CG1 CG2"},
    {"TRP", 48, "N CA C O CB CG CD1 CD2 CE2 NE1 CE3 CZ3 CH2 CZ2"},
    {"TYR", 62, "N CA C O CB CG CD1 CD2 CE1 CE2 CZ OH"},
    {"MMB", 167, "BJ"},
};

static int dfire_type(const std::string &res, const std::string &atom) {
    for (const DfireResidue &r : kDfire) {
        if (res != r.name) continue;
        std::istringstream names(r.atoms);
        std::string n;
        for (int k = 0; names >> n; k++)
            if (n == atom) return r.first + k;
        return -1;
    }
    return -1;
}

// Residues whose atoms the normal modes move (complex.cc:149-209; the
// protein, modified-protein, DNA and RNA names of complex.h).
static bool nm_residue(const std::string &r) {
    static const char *kNames[] = {"ALA", "ARG", "ASN", "ASP", "CYS", "GLU", "GLN", "GLY",
                                   "HIS", "ILE", "LEU", "LYS", "MET", "PHE", "PRO", "SER",
                                   "THR", "TRP", "TYR", "VAL", "CYX", "HIP", "HID", "HIE",
                                   "DA",  "DC",  "DI",  "DG",  "DT",  "A",   "C",   "G",
                                   "U",   "I"};
    for (const char *k : kNames)
        if (r == k) return true;
    return false;
}

// Inverse of a symmetric positive definite 4 x 4 matrix by Gauss-Jordan
// elimination with partial pivoting.
static void inverse4(const double a_in[16], double inv[16]) {
    double a[4][8];
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) {
            a[i][j] = a_in[i * 4 + j];
            a[i][4 + j] = i == j ? 1.0 : 0.0;
        }
    for (int c = 0; c < 4; c++) {
        int p = c;
        for (int r = c + 1; r < 4; r++)
            if (std::fabs(a[r][c]) > std::fabs(a[p][c])) p = r;
        if (a[p][c] == 0.0) throw std::runtime_error("singular matrix in the reference-point fit");
        if (p != c)
            for (int j = 0; j < 8; j++) std::swap(a[p][j], a[c][j]);
        const double d = a[c][c];
        for (int j = 0; j < 8; j++) a[c][j] /= d;
        for (int r = 0; r < 4; r++) {
            if (r == c) continue;
            const double f = a[r][c];
            if (f == 0.0) continue;
            for (int j = 0; j < 8; j++) a[r][j] -= f * a[c][j];
        }
    }
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) inv[i * 4 + j] = a[i][4 + j];
}

// Khachiyan's minimum-volume enclosing ellipsoid of the points, stopped when
// the weight update is below `tol` (prepare_gso_tasks.cc:106-267); returns
// the centre sum_i u_i x_i.  Only the diagonal of Q^T (Q u Q^T)^-1 Q is
// formed (the reference builds the whole N x N matrix).
static void mvee_centre(const double *xyz, int n, double tol, double centre[3]) {
    std::vector<double> u((size_t)n, 1.0 / n), nu((size_t)n);
    double err = 1.0 + tol;
    int iter = 0;
    while (err > tol) {
        if (++iter > 100000) throw std::runtime_error("reference-point fit did not converge");
        // V = Q diag(u) Q^T with q_i = (x, y, z, 1)
        double V[16] = {0};
        for (int a = 0; a < 4; a++)
            for (int b = 0; b < 4; b++) {
                double s = 0.0;
                for (int i = 0; i < n; i++) {
                    const double qa = a < 3 ? xyz[3 * (size_t)i + a] : 1.0;
                    const double qb = b < 3 ? xyz[3 * (size_t)i + b] : 1.0;
                    s += qa * (qb * u[i]);
                }
                V[a * 4 + b] = s;
            }
        double A[16];
        inverse4(V, A);
        double best = 0.0;
        int jbest = 0;
        for (int i = 0; i < n; i++) {
            const double q[4] = {xyz[3 * (size_t)i], xyz[3 * (size_t)i + 1], xyz[3 * (size_t)i + 2], 1.0};
            double m = 0.0;
            for (int k = 0; k < 4; k++) {
                double t = 0.0;
                for (int l = 0; l < 4; l++) t += A[l * 4 + k] * q[l];
                m += q[k] * t;
            }
            if (m > best) {
                best = m;
                jbest = i;
            }
        }
        const double step = (best - 3.0 - 1.0) / ((3.0 + 1.0) * (best - 1.0));
        double e2 = 0.0;
        for (int i = 0; i < n; i++) nu[i] = (1.0 - step) * u[i];
        nu[jbest] += step;
        for (int i = 0; i < n; i++) e2 += (nu[i] - u[i]) * (nu[i] - u[i]);
        err = std::sqrt(e2);
        u.swap(nu);
    }
    for (int c = 0; c < 3; c++) {
        double s = 0.0;
        for (int i = 0; i < n; i++) s += xyz[3 * (size_t)i + c] * u[i];
        centre[c] = s;
    }
}

// One atom as the reference's write_pdb formats it (prepare_gso_tasks.cc:1279-1349).
static void pdb_line(std::string &out, const AtomRec &a, int serial, double x, double y, double z) {
    auto pad_left = [](const std::string &s, size_t w) {
        return s.size() >= w ? s : std::string(w - s.size(), ' ') + s;
    };
    auto fixed = [](double v, int prec) {
        char b[64];
        std::snprintf(b, sizeof b, "%.*f", prec, v);
        return std::string(b);
    };
    std::string name = a.name.size() < 4 ? " " + a.name : a.name;
    if (name.size() < 4) name += std::string(4 - name.size(), ' ');
    const std::string alt = a.altloc.empty() ? " " : a.altloc;
    const std::string ins = a.insertion.empty() ? " " : a.insertion;
    out += a.record;
    out += pad_left(std::to_string(serial), 5);
    out += " ";
    out += name;
    out += alt;
    out += pad_left(a.resname, 3);
    out += pad_left(a.chain, 2);
    out += pad_left(std::to_string(a.resseq), 4);
    out += ins;
    out += "   ";
    out += pad_left(fixed(x, 3), 8);
    out += pad_left(fixed(y, 3), 8);
    out += pad_left(fixed(z, 3), 8);
    out += pad_left(fixed(a.occupancy, 2), 6);
    out += pad_left(fixed(a.bfactor, 2), 6);
    out += pad_left(a.element, 12);
    out += "\n";
}

// Access for the normal-mode code (anm.cpp).
int structure_atoms(sdgso_structure s) { return (int)s->atoms.size(); }
const char *structure_atom_name(sdgso_structure s, int i) { return s->atoms[i].name.c_str(); }
const char *structure_residue_name(sdgso_structure s, int r) { return s->residues[r].name.c_str(); }
void structure_residues(sdgso_structure s, std::vector<int> &first, std::vector<int> &count) {
    first.clear();
    count.clear();
    for (const Residue &r : s->residues) {
        first.push_back(r.first);
        count.push_back(r.count);
    }
}

}  // namespace sd

using namespace sd;

namespace {
template <class F>
int guard(F &&f) {
    try {
        f();
        return SDGSO_OK;
    } catch (const std::bad_alloc &) {
        set_error("host allocation failed");
        return SDGSO_ENOMEM;
    } catch (const std::exception &e) {
        set_error(e.what());
        return SDGSO_EINVAL;
    }
}
}  // namespace

int sdgso_structure_read_pdb(const char *path, int flags, sdgso_structure *out) {
    return guard([&] {
        if (!path || !out) throw std::runtime_error("null argument");
        *out = nullptr;
        std::unique_ptr<sdgso_structure_s> s(new sdgso_structure_s());
        read_pdb(path, flags & SDGSO_PDB_NO_OXT, flags & SDGSO_PDB_NO_H, flags & SDGSO_PDB_NO_WATER, *s);
        *out = s.release();
    });
}

int sdgso_structure_synthetic(int n_atoms, const double *xyz, sdgso_structure *out) {
    return guard([&] {
        if (n_atoms <= 0 || !xyz || !out) throw std::runtime_error("bad arguments");
        std::unique_ptr<sdgso_structure_s> s(new sdgso_structure_s());
        s->chains.push_back("A");
        for (int i = 0; i < n_atoms; i++) {
            AtomRec a;
            a.record = "ATOM  ";
            a.serial = i + 1;
            a.name = "CA";
            a.resname = "ALA";
            a.chain = "A";
            a.insertion = " ";
            a.resseq = i + 1;
            a.element = "C";
            a.x = xyz[3 * (size_t)i];
            a.y = xyz[3 * (size_t)i + 1];
            a.z = xyz[3 * (size_t)i + 2];
            Residue r;
            r.name = "ALA";
            r.number = i + 1;
            r.insertion = " ";
            r.first = i;
            r.count = 1;
            s->residues.push_back(r);
            s->atoms.push_back(a);
        }
        *out = s.release();
    });
}

int sdgso_structure_destroy(sdgso_structure s) {
    delete s;
    return SDGSO_OK;
}

int sdgso_structure_size(sdgso_structure s, int *n_atoms, int *n_residues, int *n_chains) {
    return guard([&] {
        if (!s) throw std::runtime_error("null structure");
        if (n_atoms) *n_atoms = (int)s->atoms.size();
        if (n_residues) *n_residues = (int)s->residues.size();
        if (n_chains) *n_chains = (int)s->chains.size();
    });
}

int sdgso_structure_coords(sdgso_structure s, double *xyz) {
    return guard([&] {
        if (!s || !xyz) throw std::runtime_error("null argument");
        for (size_t i = 0; i < s->atoms.size(); i++) {
            xyz[3 * i] = s->atoms[i].x;
            xyz[3 * i + 1] = s->atoms[i].y;
            xyz[3 * i + 2] = s->atoms[i].z;
        }
    });
}

int sdgso_structure_atom(sdgso_structure s, int i, char *name, char *resname, int *resseq,
                         char *chain) {
    return guard([&] {
        if (!s || i < 0 || i >= (int)s->atoms.size()) throw std::runtime_error("bad atom index");
        const AtomRec &a = s->atoms[i];
        if (name) std::snprintf(name, 8, "%s", a.name.c_str());
        if (resname) std::snprintf(resname, 8, "%s", a.resname.c_str());
        if (resseq) *resseq = a.resseq;
        if (chain) std::snprintf(chain, 4, "%s", a.chain.c_str());
    });
}

int sdgso_structure_dfire_types(sdgso_structure s, int *types) {
    return guard([&] {
        if (!s || !types) throw std::runtime_error("null argument");
        for (const Residue &r : s->residues)
            for (int i = r.first; i < r.first + r.count; i++) {
                const int t = dfire_type(r.name, s->atoms[i].name);
                if (t < 0)
                    throw std::runtime_error("no DFIRE type for atom " + std::to_string(s->atoms[i].serial) +
                                             " (" + r.name + " " + s->atoms[i].name + ")");
                types[i] = t;
            }
    });
}

int sdgso_structure_nm_mask(sdgso_structure s, int *mask) {
    return guard([&] {
        if (!s || !mask) throw std::runtime_error("null argument");
        for (const Residue &r : s->residues)
            for (int i = r.first; i < r.first + r.count; i++) mask[i] = nm_residue(r.name) ? 1 : 0;
    });
}

int sdgso_structure_backbone(sdgso_structure s, int *index, int *n) {
    return guard([&] {
        if (!s || !n) throw std::runtime_error("null argument");
        int k = 0;
        for (size_t i = 0; i < s->atoms.size(); i++)
            if (s->atoms[i].name == "CA" || s->atoms[i].name == "P") {
                if (index) index[k] = (int)i;
                k++;
            }
        *n = k;
    });
}

int sdgso_structure_centre(sdgso_structure s, double *translation) {
    return guard([&] {
        if (!s || !translation) throw std::runtime_error("null argument");
        double t[3] = {0, 0, 0};
        int n = 0;
        for (const AtomRec &a : s->atoms)
            if (a.element != "H") {
                t[0] += a.x;
                t[1] += a.y;
                t[2] += a.z;
                n++;
            }
        if (n == 0) throw std::runtime_error("no heavy atoms");
        for (int c = 0; c < 3; c++) translation[c] = -1 * t[c] / n;
    });
}

int sdgso_reference_point(int n_atoms, const double *xyz, double tolerance, double *centre) {
    return guard([&] {
        if (n_atoms <= 0 || !xyz || !centre || !(tolerance > 0)) throw std::runtime_error("bad arguments");
        mvee_centre(xyz, n_atoms, tolerance, centre);
    });
}

int sdgso_structure_write_pdb(sdgso_structure s, const double *xyz, const char *path, int append) {
    return guard([&] {
        if (!s || !path) throw std::runtime_error("null argument");
        std::string out;
        out.reserve(s->atoms.size() * 82);
        for (size_t i = 0; i < s->atoms.size(); i++) {
            const AtomRec &a = s->atoms[i];
            pdb_line(out, a, (int)i + 1, xyz ? xyz[3 * i] : a.x, xyz ? xyz[3 * i + 1] : a.y,
                     xyz ? xyz[3 * i + 2] : a.z);
        }
        FILE *f = std::fopen(path, append ? "ab" : "wb");
        if (!f) throw std::runtime_error(std::string("cannot write ") + path);
        const size_t w = std::fwrite(out.data(), 1, out.size(), f);
        std::fclose(f);
        if (w != out.size()) throw std::runtime_error(std::string("short write to ") + path);
    });
}
