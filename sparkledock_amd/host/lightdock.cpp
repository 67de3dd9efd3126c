// lightdock: the command-line driver of the GSO run, with the reference's
// interface `lightdock -f setup.json -s STEPS [-l N]`
// (source/simulation/run_simulation.cc:21-130).
//
// Differences from the reference driver, by design:
//   - `-s STEPS` is honoured (the reference ignores it, run_simulation.cc:37-38);
//     `-l` is accepted and ignored, as there;
//   - the GSO keys of setup.json (rho, gamma, beta, initialVisionRange,
//     initialLuciferin, maximumNeighbors, maximumVisionRange) are honoured when
//     present (read_input_structure.cc:49-55 reads them but never uses them);
//   - one GPU runs all swarms: no MPI;
//   - the run always executes on the GPU through libsdgso (USE_CUDA is not
//     consulted); "form": "fp64" selects the plain FP64 kernels, otherwise the
//     FP32 guard-band form runs;
//   - inputs are either the reference's "receptor_pdb" / "ligand_pdb" (read
//     with its noxt / noh / now filters, DFIRE-typed, masked, centred, normal
//     modes from the ANM when use_anm is set, the ligand reference point from
//     the minimum-volume ellipsoid, starting poses from the reference's
//     surface sampling with "swarms" given, as this build does not link
//     FreeSASA to derive it) or synthetic complexes ("synthetic_receptor_atoms",
//     "synthetic_ligand_atoms", "synthetic_seed", seeded placement around
//     the receptor); the DFIRE table is read from "dfire_table" (one value
//     per line, the DCparams layout, kept in FP64) or drawn from the
//     synthetic N(0,1)*0.5 table;
//   - every swarm_<i>/gso_<STEPS>.out is written fresh (the reference's mmap
//     writer appends to stale files, prepare_gso_task.cu:2860-2904);
//   - swarm_<i>/cluster.repr is written after the run (the reference's
//     clustering code, prepare_gso_tasks.cc:961-1120, sits after its exit(0));
//     "write_pdbs": true also writes swarm_<i>/lightdock_<g>.pdb.
// Output format of each line follows prepare_gso_task.cu:2868-2888.
#include <getopt.h>
#include <sys/stat.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "sdgso.h"

namespace {

// Minimal reader of a flat JSON object: string, number, boolean and null values.
std::map<std::string, std::string> read_flat_json(const std::string &path) {
    std::ifstream in(path);
    if (!in) throw std::runtime_error("cannot open " + path);
    std::stringstream ss;
    ss << in.rdbuf();
    const std::string s = ss.str();
    std::map<std::string, std::string> kv;
    size_t i = s.find('{');
    if (i == std::string::npos) throw std::runtime_error("setup file is not a JSON object");
    i++;
    auto skip = [&] {
        while (i < s.size() && (std::isspace((unsigned char)s[i]) || s[i] == ',')) i++;
    };
    auto str = [&]() {
        std::string out;
        if (s[i] != '"') throw std::runtime_error("expected a string in setup file");
        i++;
        while (i < s.size() && s[i] != '"') {
            if (s[i] == '\\' && i + 1 < s.size()) i++;
            out += s[i++];
        }
        i++;
        return out;
    };
    for (;;) {
        skip();
        if (i >= s.size() || s[i] == '}') break;
        const std::string key = str();
        skip();
        if (s[i] != ':') throw std::runtime_error("expected ':' after " + key);
        i++;
        skip();
        std::string val;
        if (s[i] == '"') {
            val = str();
        } else if (s[i] == '{' || s[i] == '[') {
            int depth = 0;  // nested values are skipped (kept as raw text)
            do {
                if (s[i] == '{' || s[i] == '[') depth++;
                if (s[i] == '}' || s[i] == ']') depth--;
                val += s[i++];
            } while (i < s.size() && depth > 0);
        } else {
            while (i < s.size() && s[i] != ',' && s[i] != '}' && !std::isspace((unsigned char)s[i]))
                val += s[i++];
        }
        kv[key] = val;
    }
    return kv;
}

struct Setup {
    std::map<std::string, std::string> kv;
    double num(const char *k, double def) const {
        auto it = kv.find(k);
        if (it == kv.end() || it->second == "null") return def;
        if (it->second == "true") return 1;
        if (it->second == "false") return 0;
        return std::strtod(it->second.c_str(), nullptr);
    }
    std::string text(const char *k, const std::string &def) const {
        auto it = kv.find(k);
        return it == kv.end() ? def : it->second;
    }
};

void check(int rc, const char *what) {
    if (rc != SDGSO_OK) {
        std::fprintf(stderr, "%s failed (%d): %s\n", what, rc, sdgso_last_error());
        std::exit(1);
    }
}

// std::to_string(double) formatting ("%f"), as the reference writer uses.
std::string f6(double v) { return std::to_string(v); }

// swarm_<s>/cluster.repr: "cluster:size:score:representative:lightdock_<rep>.pdb"
// per cluster in order of foundation (prepare_gso_tasks.cc:1084-1110; the
// score printed is the representative's own, and every member is counted once).
void write_clusters(int s, int G, const int *cl, const int *repr, const std::vector<double> &score) {
    const std::string path = "swarm_" + std::to_string(s) + "/cluster.repr";
    std::vector<int> size(G, 0), first(G, -1);
    int nc = 0;
    for (int g = 0; g < G; g++) {
        size[cl[g]]++;
        first[cl[g]] = repr[g];
        nc = std::max(nc, cl[g] + 1);
    }
    std::string buf;
    char line[160];
    for (int c = 0; c < nc; c++) {
        const int id = first[c];
        std::snprintf(line, sizeof line, "%d:%d:%g:%d:lightdock_%d.pdb\n", c, size[c],
                      score[(size_t)s * G + id], id, id);
        buf += line;
    }
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f) {
        std::perror(path.c_str());
        std::exit(1);
    }
    std::fwrite(buf.data(), 1, buf.size(), f);
    std::fclose(f);
}

std::string resolve(const std::string &p, const std::string &setup_path) {
    struct stat sb;
    if (stat(p.c_str(), &sb) == 0 || p.empty() || p[0] == '/') return p;
    const size_t slash = setup_path.rfind('/');
    if (slash == std::string::npos) return p;
    return setup_path.substr(0, slash + 1) + p;
}

// Reads a structure and the arrays the run needs from it.
struct Molecule {
    sdgso_structure s = nullptr;
    int n = 0;
    std::vector<double> raw, xyz;  // as read; centred
    std::vector<int> type, mask;
    double translation[3] = {0, 0, 0};
};

Molecule read_molecule(const std::string &path, int flags) {
    Molecule m;
    check(sdgso_structure_read_pdb(path.c_str(), flags, &m.s), path.c_str());
    check(sdgso_structure_size(m.s, &m.n, nullptr, nullptr), "structure size");
    m.xyz.resize(3 * (size_t)m.n);
    m.type.resize(m.n);
    m.mask.resize(m.n);
    check(sdgso_structure_coords(m.s, m.xyz.data()), "coordinates");
    check(sdgso_structure_dfire_types(m.s, m.type.data()), "DFIRE typing");
    check(sdgso_structure_nm_mask(m.s, m.mask.data()), "normal-mode mask");
    check(sdgso_structure_centre(m.s, m.translation), "centring");
    m.raw = m.xyz;
    for (int i = 0; i < m.n; i++)
        for (int k = 0; k < 3; k++) m.xyz[3 * (size_t)i + k] += m.translation[k];
    return m;
}

void write_swarm(int s, int steps, int G, int P, const std::vector<double> &pos,
                 const std::vector<double> &lucif, const std::vector<int> &nnei,
                 const std::vector<double> &vis, const std::vector<double> &score) {
    const std::string dir = "swarm_" + std::to_string(s);
    mkdir(dir.c_str(), 0777);
    const std::string path = dir + "/gso_" + std::to_string(steps) + ".out";
    std::string buf = "#Coordinates  RecID  LigID  Luciferin  Neighbor's number  Vision Range  Scoring\n";
    for (int g = 0; g < G; g++) {
        const size_t id = (size_t)s * G + g;
        std::string line = "(";
        for (int p = 0; p < P; p++) {
            line += f6(pos[id * P + p]);
            if (p < P - 1) line += ", ";
        }
        line += ")\t" + std::to_string(0) + "\t" + std::to_string(0) + "\t" + f6(lucif[id]) + "\t" +
                std::to_string(nnei[id]) + "\t" + f6(vis[id]) + "\t" + f6(score[id]) + "\t\n";
        buf += line;
    }
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f) {
        std::perror(path.c_str());
        std::exit(1);
    }
    std::fwrite(buf.data(), 1, buf.size(), f);
    std::fclose(f);
}

}  // namespace

int main(int argc, char **argv) {
    std::string setup_path;
    int steps = -1;
    int ch;
    while ((ch = getopt(argc, argv, "f:s:l:")) != -1) {
        if (ch == 'f') setup_path = optarg;
        else if (ch == 's') steps = std::atoi(optarg);
        else if (ch == 'l') { /* accepted and ignored, as in the reference */ }
    }
    if (setup_path.empty() || steps < 0) {
        std::fprintf(stderr, "usage: lightdock -f setup.json -s STEPS [-l N]\n");
        return 2;
    }
    Setup st;
    try {
        st.kv = read_flat_json(setup_path);
    } catch (const std::exception &e) {
        std::fprintf(stderr, "%s\n", e.what());
        return 1;
    }
    const std::string rec_pdb = st.text("receptor_pdb", ""), lig_pdb = st.text("ligand_pdb", "");
    const bool from_pdb = !rec_pdb.empty() && !lig_pdb.empty() && rec_pdb != "null" && lig_pdb != "null";
    const uint64_t seed = (uint64_t)st.num("synthetic_seed", 0);
    const bool use_anm = st.num("use_anm", 0) != 0;
    int nmr = use_anm ? (int)st.num("anm_rec", 10) : 0;
    int nml = use_anm ? (int)st.num("anm_lig", 10) : 0;
    const int S = (int)st.num("swarms", 1);
    const int G = (int)st.num("glowworms", 200);
    const double swarm_radius = st.num("swarm_radius", 10.0);
    const int form = st.text("form", "fp32_guard") == "fp64" ? SDGSO_FP64 : SDGSO_FP32_GUARD;
    sdgso_params prm;
    sdgso_default_params(&prm);
    prm.rho = st.num("rho", prm.rho);
    prm.gamma = st.num("gamma", prm.gamma);
    prm.beta = st.num("beta", prm.beta);
    prm.vision0 = st.num("initialVisionRange", prm.vision0);
    prm.luciferin0 = st.num("initialLuciferin", prm.luciferin0);
    prm.max_neighbors = (int)st.num("maximumNeighbors", prm.max_neighbors);
    prm.max_vision = st.num("maximumVisionRange", prm.max_vision);
    prm.step_translation = st.num("translation_step", prm.step_translation);
    prm.step_rotation = st.num("rotation_step", prm.step_rotation);
    prm.step_nmodes = st.num("nmodes_step", prm.step_nmodes);

    int nr = 0, nl = 0;
    std::vector<double> rx, lx, rm(1), lm(1), pos;
    std::vector<int> rt, lt, rmask, lmask;
    sdgso_structure rs = nullptr, ls = nullptr;
    double ref[3] = {0, 0, 0};
    if (from_pdb) {
        const int flags = (st.num("noxt", 0) != 0 ? SDGSO_PDB_NO_OXT : 0) |
                          (st.num("noh", 0) != 0 ? SDGSO_PDB_NO_H : 0) |
                          (st.num("now", 0) != 0 ? SDGSO_PDB_NO_WATER : 0);
        Molecule rmol = read_molecule(resolve(rec_pdb, setup_path), flags);
        Molecule lmol = read_molecule(resolve(lig_pdb, setup_path), flags);
        nr = rmol.n;
        nl = lmol.n;
        rx = rmol.xyz;
        lx = lmol.xyz;
        rt = rmol.type;
        lt = lmol.type;
        rmask = rmol.mask;
        lmask = lmol.mask;
        rs = rmol.s;
        ls = lmol.s;
        // starting poses from the coordinates as read (run_simulation.cc:59-71)
        pos.resize((size_t)S * G * (7 + nmr + nml));
        check(sdgso_starting_poses(rs, rmol.raw.data(), nl, lmol.raw.data(), rmol.translation, S, G,
                                   swarm_radius, (uint32_t)st.num("starting_points_seed", 324324),
                                   (uint32_t)st.num("anm_seed", 324324), nmr, nml, nullptr,
                                   pos.data()),
              "starting poses");
        // normal modes (cal_anm.cc:61-219, run_simulation.cc:79-92); the coordinates
        // of a molecule with modes are re-centred on all its atoms
        if (nmr > 0) {
            rm.assign((size_t)nmr * nr * 3, 0.0);
            check(sdgso_structure_anm(rs, rx.data(), nmr, st.num("anm_rec_rmsd", 0.5), S * G,
                                      rm.data(), nullptr),
                  "receptor normal modes");
        }
        if (nml > 0) {
            lm.assign((size_t)nml * nl * 3, 0.0);
            check(sdgso_structure_anm(ls, lx.data(), nml, st.num("anm_lig_rmsd", 0.5), S * G,
                                      lm.data(), nullptr),
                  "ligand normal modes");
        }
        check(sdgso_reference_point(nl, lx.data(), 0.01, ref), "reference point");
    } else {
        nr = (int)st.num("synthetic_receptor_atoms", 0);
        nl = (int)st.num("synthetic_ligand_atoms", 0);
        if (nr <= 0 || nl <= 0) {
            std::fprintf(stderr,
                         "no input complex: set receptor_pdb and ligand_pdb, or "
                         "synthetic_receptor_atoms and synthetic_ligand_atoms, in %s\n",
                         setup_path.c_str());
            return 1;
        }
        rx.resize(3 * (size_t)nr);
        lx.resize(3 * (size_t)nl);
        rm.resize((size_t)nmr * nr * 3 + 1);
        lm.resize((size_t)nml * nl * 3 + 1);
        rt.resize(nr);
        lt.resize(nl);
        check(sdgso_synth_molecule(1000 * seed + 1, nr, nmr, rx.data(), rt.data(), rm.data()), "synth");
        check(sdgso_synth_molecule(1000 * seed + 2, nl, nml, lx.data(), lt.data(), lm.data()), "synth");
        for (int i = 0; i < nl; i++)
            for (int k = 0; k < 3; k++) ref[k] += lx[3 * (size_t)i + k] / nl;
        check(sdgso_structure_synthetic(nr, rx.data(), &rs), "structure");
        check(sdgso_structure_synthetic(nl, lx.data(), &ls), "structure");
    }
    // FP64, as FastDifire::difire_energy holds it (fastdfire.h:93)
    std::vector<double> table(SDGSO_TABLE_LEN, 0.0);
    const std::string table_path = st.text("dfire_table", "");
    if (!table_path.empty() && table_path != "null") {
        std::ifstream tin(resolve(table_path, setup_path));
        if (!tin) {
            std::fprintf(stderr, "cannot open the DFIRE table %s\n", table_path.c_str());
            return 1;
        }
        double v;
        int n = 0;
        while (n < SDGSO_NTYPES * SDGSO_NTYPES * SDGSO_NBINS && (tin >> v)) table[n++] = v;
        if (n != SDGSO_NTYPES * SDGSO_NTYPES * SDGSO_NBINS) {
            std::fprintf(stderr, "the DFIRE table %s holds %d values, expected %d\n",
                         table_path.c_str(), n, SDGSO_NTYPES * SDGSO_NTYPES * SDGSO_NBINS);
            return 1;
        }
    } else {
        std::vector<float> t32(SDGSO_TABLE_LEN);
        check(sdgso_synth_table(7, t32.data()), "synth table");
        for (int i = 0; i < SDGSO_TABLE_LEN; i++) table[i] = t32[i];
    }
    const int P = 7 + nmr + nml;
    if (!from_pdb) {
        // synthetic complexes: swarm centres on a sphere around the globule
        pos.resize((size_t)S * G * P);
        const double centre = std::cbrt(nr * 11.4) + 0.5 * std::cbrt(nl * 11.4);
        check(sdgso_synth_poses((uint64_t)st.num("starting_points_seed", 1000 * seed + 3), S, G, nmr,
                                nml, centre, swarm_radius, pos.data()),
              "starting poses");
    }
    // the reference's stream: mt19937(0), [swarm][step][glowworm]
    std::vector<double> u((size_t)S * steps * G + 1);
    check(sdgso_stream(0, S, steps, G, u.data()), "stream");

    sdgso_complex cx = nullptr;
    check(sdgso_complex_create(nr, rx.data(), rt.data(), rmask.empty() ? nullptr : rmask.data(), nmr,
                               nmr ? rm.data() : nullptr, nl, lx.data(), lt.data(),
                               lmask.empty() ? nullptr : lmask.data(), nml,
                               nml ? lm.data() : nullptr, ref, table.data(), &cx),
          "sdgso_complex_create");
    sdgso_run run = nullptr;
    check(sdgso_run_create(1, &cx, &S, G, pos.data(), u.data(), steps, &prm, form, 1e-4, &run),
          "sdgso_run_create");
    if (form == SDGSO_FP32_GUARD) {
        // pick the FP32 scoring kernel for this complex (scores the starting
        // poses a few times; the run's state is unchanged)
        int chosen = 0;
        check(sdgso_run_tune(run, 1, nullptr, &chosen), "sdgso_run_tune");
        int chunk = 0, lgemm = 0;
        check(sdgso_run_work_order(run, &chunk, nullptr), "sdgso_run_work_order");
        check(sdgso_run_ligand_gemm(run, &lgemm, nullptr, nullptr), "sdgso_run_ligand_gemm");
        std::printf("scoring kernel: %s, work items %s, ligand displacement %s\n",
                    sdgso_tune_variant_name(chosen), chunk ? "in chunks per XCD" : "in list order",
                    lgemm ? "by the matrix-core GEMM" : "in the scoring set-up");
    }
    std::printf("prepare docking\n");
    const auto t0 = std::chrono::steady_clock::now();
    check(sdgso_run_steps(run, steps, 1), "sdgso_run_steps");
    check(sdgso_run_sync(run), "sdgso_run_sync");
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    const size_t N = (size_t)S * G;
    std::vector<double> lucif(N), vis(N), score(N);
    std::vector<int> nnei(N);
    check(sdgso_run_read(run, pos.data(), lucif.data(), score.data(), vis.data(), nnei.data(),
                         nullptr, nullptr, nullptr, nullptr),
          "sdgso_run_read");
    std::printf("saving pos file\n");
    for (int s = 0; s < S; s++) write_swarm(s, steps, G, P, pos, lucif, nnei, vis, score);
    // clustering over the CA / P atoms of both molecules, 4 A (prepare_gso_tasks.cc:961-1080)
    int nrb = 0, nlb = 0;
    check(sdgso_structure_backbone(rs, nullptr, &nrb), "backbone");
    check(sdgso_structure_backbone(ls, nullptr, &nlb), "backbone");
    std::vector<int> rb(nrb + 1), lb(nlb + 1), cl(N), repr(N);
    check(sdgso_structure_backbone(rs, rb.data(), &nrb), "backbone");
    check(sdgso_structure_backbone(ls, lb.data(), &nlb), "backbone");
    check(sdgso_run_cluster(run, 0, rb.data(), nrb, lb.data(), nlb, 4.0, cl.data(), repr.data()),
          "sdgso_run_cluster");
    for (int s = 0; s < S; s++) write_clusters(s, G, cl.data() + (size_t)s * G, repr.data() + (size_t)s * G, score);
    if (st.num("write_pdbs", 0) != 0) {
        std::vector<double> rc((size_t)G * nr * 3), lc((size_t)G * nl * 3);
        for (int s = 0; s < S; s++) {
            check(sdgso_run_final_coords(run, s * G, G, rc.data(), lc.data()), "final coordinates");
            for (int g = 0; g < G; g++) {
                const std::string path = "swarm_" + std::to_string(s) + "/lightdock_" + std::to_string(g) + ".pdb";
                check(sdgso_structure_write_pdb(rs, rc.data() + (size_t)g * nr * 3, path.c_str(), 0), "write_pdb");
                check(sdgso_structure_write_pdb(ls, lc.data() + (size_t)g * nl * 3, path.c_str(), 1), "write_pdb");
            }
        }
    }
    std::printf("simulation execution:%f\n", secs);
    sdgso_run_destroy(run);
    sdgso_complex_destroy(cx);
    sdgso_structure_destroy(rs);
    sdgso_structure_destroy(ls);
    return 0;
}
