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
//   - inputs are synthetic complexes ("synthetic_receptor_atoms",
//     "synthetic_ligand_atoms", "synthetic_seed"); PDB input is not read by
//     this build;
//   - every swarm_<i>/gso_<STEPS>.out is written fresh (the reference's mmap
//     writer appends to stale files, prepare_gso_task.cu:2860-2904).
// Output format of each line follows prepare_gso_task.cu:2868-2888.
#include <getopt.h>
#include <sys/stat.h>

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
    const int nr = (int)st.num("synthetic_receptor_atoms", 0);
    const int nl = (int)st.num("synthetic_ligand_atoms", 0);
    if (nr <= 0 || nl <= 0) {
        std::fprintf(stderr,
                     "this build docks synthetic complexes: set synthetic_receptor_atoms and "
                     "synthetic_ligand_atoms in %s (PDB input is not read)\n",
                     setup_path.c_str());
        return 1;
    }
    const uint64_t seed = (uint64_t)st.num("synthetic_seed", 0);
    const bool use_anm = st.num("use_anm", 0) != 0;
    const int nmr = use_anm ? (int)st.num("anm_rec", 10) : 0;
    const int nml = use_anm ? (int)st.num("anm_lig", 10) : 0;
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

    std::vector<double> rx(3 * (size_t)nr), lx(3 * (size_t)nl), rm((size_t)nmr * nr * 3 + 1),
        lm((size_t)nml * nl * 3 + 1);
    std::vector<int> rt(nr), lt(nl);
    std::vector<float> table(SDGSO_TABLE_LEN);
    check(sdgso_synth_molecule(1000 * seed + 1, nr, nmr, rx.data(), rt.data(), rm.data()), "synth");
    check(sdgso_synth_molecule(1000 * seed + 2, nl, nml, lx.data(), lt.data(), lm.data()), "synth");
    check(sdgso_synth_table(7, table.data()), "synth table");
    double ref[3] = {0, 0, 0};
    for (int i = 0; i < nl; i++)
        for (int k = 0; k < 3; k++) ref[k] += lx[3 * (size_t)i + k] / nl;
    const int P = 7 + nmr + nml;
    std::vector<double> pos((size_t)S * G * P);
    const double centre = std::cbrt(nr * 11.4) + 0.5 * std::cbrt(nl * 11.4);
    check(sdgso_synth_poses((uint64_t)st.num("starting_points_seed", 1000 * seed + 3), S, G, nmr,
                            nml, centre, swarm_radius, pos.data()),
          "synth poses");
    // the reference's stream: mt19937(0), [swarm][step][glowworm]
    std::vector<double> u((size_t)S * steps * G + 1);
    check(sdgso_stream(0, S, steps, G, u.data()), "stream");

    sdgso_complex cx = nullptr;
    check(sdgso_complex_create(nr, rx.data(), rt.data(), nullptr, nmr, nmr ? rm.data() : nullptr, nl,
                               lx.data(), lt.data(), nullptr, nml, nml ? lm.data() : nullptr, ref,
                               table.data(), &cx),
          "sdgso_complex_create");
    sdgso_run run = nullptr;
    check(sdgso_run_create(1, &cx, &S, G, pos.data(), u.data(), steps, &prm, form, 1e-4, &run),
          "sdgso_run_create");
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
    std::printf("simulation execution:%f\n", secs);
    sdgso_run_destroy(run);
    sdgso_complex_destroy(cx);
    return 0;
}
