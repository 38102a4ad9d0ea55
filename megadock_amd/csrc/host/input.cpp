// Input layer of the drop-in: command line, tables, rotation sets, PDB reader, per-atom
// parameters and grid sizing (src/parameter.cpp, src/pdb_entry.cpp, src/protein.cpp,
// src/control.cpp:91-465).
#pragma clang fp contract(off)

#include <dlfcn.h>
#include <getopt.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>

#include "mdk_host.h"

namespace mdk {

std::string data_dir() {
    if (const char *e = std::getenv("MDK_DATA_DIR")) return e;
    Dl_info info;
    if (dladdr((void *)&data_dir, &info) && info.dli_fname) {
        std::string p = info.dli_fname;  // .../megadock_amd/lib/libmdk.so or .../bin/megadock-hip
        const size_t s = p.rfind('/');
        std::string dir = s == std::string::npos ? "." : p.substr(0, s);
        return dir + "/../data";
    }
    return "megadock_amd/data";
}

// ------------------------------------------------------------------ Parameter
std::string Parameter::usage() {
    return "\n Usage: megadock-hip -R [receptor.pdb] -L [ligand.pdb] [more options]\n"
           "\n Available options:\n"
           "  -o filename    : set the output filename (default to \"$R-$L.out\")\n"
           "  -O             : output docking detail files (not produced by this build)\n"
           "  -N integer     : set the number of output predictions (default to 2000)\n"
           "  -t integer     : set the number of predictions per each rotation (default to 1, at most 16)\n"
           "  -F integer     : set the number of FFT point (default to none)\n"
           "  -v float       : set the voxel size (default to 1.2)\n"
           "  -D             : set the 6 deg. (54000 angles) of rotational sampling\n"
           "  -r integer     : set the number of rotational sampling angles (54000, 3600, 360, 24, 3, 1)\n"
           "  -e float       : set the electrostatics term ratio (default to 1.0)\n"
           "  -d float       : set the hydrophobic term ratio (default to 1.0)\n"
           "  -a float       : set the rPSC receptor core penalty (default to -45.0)\n"
           "  -b float       : set the rPSC ligand core penalty (default to 1.0)\n"
           "  -f 1/2/3       : set function (1: rPSC, 2: rPSC+Elec, 3: rPSC+Elec+RDE, default to 3)\n"
           "  -h             : show this message\n";
}

static std::string basename_noext(std::string f) {  // src/parameter.cpp:320-352
    const size_t s = f.rfind('/');
    if (s != std::string::npos) f = f.substr(s + 1);
    const size_t d = f.rfind('.');
    if (d != std::string::npos) f = f.substr(0, d);
    return f;
}

int Parameter::initialize(int argc, char **argv, std::string &err) {
    bool have_r = false, have_l = false, have_o = false;
    int f1 = 0, f2 = 0;
    // optind = 0 makes glibc's getopt start afresh: with 1 it may resume inside the last
    // argument of an earlier parse (an argument vector that may no longer exist)
    optind = 0;
    int ch;
    while ((ch = getopt(argc, argv, "R:L:B:U:o:ON:t:E:a:b:e:d:F:v:Dhy:z:r:f:i:jkl:S:G:T:W:")) != -1) {
        switch (ch) {
            case 'R': rec_file = optarg; have_r = true; break;
            case 'L': lig_file = optarg; have_l = true; break;
            case 'B':
                rec_file = std::string(optarg) + "_r_b.pdb";
                lig_file = std::string(optarg) + "_l_b.pdb";
                have_r = have_l = true;
                break;
            case 'U':
                rec_file = std::string(optarg) + "_r_u.pdb";
                lig_file = std::string(optarg) + "_l_u.pdb";
                have_r = have_l = true;
                break;
            case 'o': {  // src/parameter.cpp:81-92
                out_file = optarg;
                have_o = true;
                const bool dot_out = out_file.size() > 4 && out_file.compare(out_file.size() - 4, 4, ".out") == 0;
                const std::string stem = dot_out ? out_file.substr(0, out_file.size() - 4) : out_file;
                detail_file = stem + ".detail";
                csv_file = stem + ".csv";
                break;
            }
            case 'O': detail_output_flag = 1; break;
            case 'N': num_output = std::atoi(optarg); std::printf("# Number of output = %d\n", num_output); break;
            case 't': num_sort = std::atoi(optarg); std::printf("# Set number of scores per one angle = %d\n", num_sort); break;
            case 'E': case 'i': break;  // calculation-time log: not produced
            case 'e':
                elec_ratio = (float)std::atof(optarg);
                if (f1 == 1) { err = "Do not use -f 1 and -e option at the same time."; return 1; }
                f1 = -1;
                break;
            case 'd':
                ace_ratio = (float)std::atof(optarg);
                if (f1 == 1 || f2 == 1) { err = "Do not use -f 1 or 2 and -d option at the same time."; return 1; }
                f1 = -1;
                f2 = -1;
                break;
            case 'a': rec_core = (float)std::atof(optarg); break;
            case 'b': lig_core = (float)std::atof(optarg); break;
            case 'F':
                num_fft = std::atoi(optarg);
                num_grid = num_fft / 2;
                num_fft_flag = 1;
                std::printf("# Number of FFT N  = %d\n", num_fft);
                break;
            case 'v': grid_width = (float)std::atof(optarg); break;
            case 'T':
                thread_limit = std::atoi(optarg);
                if (thread_limit < 1) { err = "Please set a positive number (Number of cores)."; return 1; }
                break;
            case 'G':
                gpu_limit = std::atoi(optarg);
                if (gpu_limit < 0) { err = "Please set a non-negative number (Number of GPUs you want to use)."; return 1; }
                break;
            case 'f':
                score_func = std::atoi(optarg);
                if (score_func < 1 || score_func > 3) { err = "score function must be 1, 2 or 3"; return 1; }
                if (score_func == 1) {
                    elec_ratio = 0.0f;
                    ace_ratio = 0.0f;
                    if (f1 == -1) { err = "Do not use -f 1 and -e/-d option at the same time."; return 1; }
                    f1 = 1;
                } else if (score_func == 2) {
                    ace_ratio = 0.0f;
                    if (f2 == -1) { err = "Do not use -f 2 and -d option at the same time."; return 1; }
                    f2 = 1;
                }
                score_func = 3;  // src/parameter.cpp:188
                break;
            case 'r': rotation_set = std::atoi(optarg); break;
            case 'D': rotation_set = 54000; break;
            case 'j': progress = 1; break;
            case 'k': load_balance = 1; break;
            case 'S': break;  // serial GPU charge kernel: one semantics here (the CPU's)
            case 'l': fft_base_set = std::atoi(optarg); break;
            case 'W': batch = std::atoi(optarg); break;  // rotations per device batch
            case 'h':
            default: err = usage(); return 2;
        }
    }
    if (!have_r) { err = "[ERROR] Receptor PDB file is not assignment!!"; return 1; }
    if (!have_l) { err = "[ERROR] Ligand PDB file is not assignment!!"; return 1; }
    if (!have_o) {  // src/parameter.cpp:320-365
        const std::string stem = basename_noext(rec_file) + "-" + basename_noext(lig_file);
        out_file = stem + ".out";
        detail_file = stem + ".detail";
        csv_file = stem + ".csv";
    }
    std::printf("# Output file = %s\n", out_file.c_str());
    if (num_sort < 1 || num_sort > 1048576) { err = "-t must be between 1 and 1048576"; return 1; }
    if (load_rotation_set(rotation_set, zangle, err)) return 1;
    num_rot_angles = (int)(zangle.size() / 3);
    if (num_rot_angles * num_sort < num_output) {  // src/parameter.cpp:396-400
        std::fprintf(stderr, "#[WARNING] Number of output file was changed from %d to %d\n", num_output,
                     num_rot_angles * num_sort);
        num_output = num_rot_angles * num_sort;
    }
    return load_tables(err);
}

int Parameter::load_rotation_set(int n, std::vector<float> &out, std::string &err) {
    if (n != 1 && n != 3 && n != 24 && n != 360 && n != 3600 && n != 54000) {
        err = "[ERROR] Not support Rotation_angle_mode = " + std::to_string(n);
        return 1;
    }
    const std::string path = data_dir() + "/rotations_" + std::to_string(n) + ".txt";
    std::ifstream f(path);
    if (!f) { err = "cannot open " + path; return 1; }
    out.clear();
    std::string line;
    while (std::getline(f, line)) {
        if (line.empty() || line[0] == '#') continue;
        std::istringstream is(line);
        std::string a;
        while (is >> a) out.push_back((float)std::strtod(a.c_str(), nullptr));  // double literal -> float
    }
    if ((int)out.size() != 3 * n) { err = "bad rotation table " + path; return 1; }
    return 0;
}

static int load_map(const std::string &path, std::map<std::string, float> &m, std::string &err) {
    std::ifstream f(path);
    if (!f) { err = "cannot open " + path; return 1; }
    std::string line;
    while (std::getline(f, line)) {
        if (line.empty() || line[0] == '#') continue;
        const size_t tab = line.find('\t');
        if (tab == std::string::npos) continue;
        m.insert({line.substr(0, tab), (float)std::strtod(line.c_str() + tab + 1, nullptr)});  // first wins
    }
    return 0;
}

int Parameter::load_tables(std::string &err) {
    const std::string d = data_dir();
    if (load_map(d + "/charmm_radius.tsv", charmm_r_, err) || load_map(d + "/charmm_charge.tsv", charmm_c_, err) ||
        load_map(d + "/ace.tsv", ace_, err))
        return 1;
    return 0;
}

// src/parameter.cpp:425-470: missing keys fall back to radius 1.8, charge 0, ACE 0
float Parameter::atom_radius(const std::string &k) const {
    auto it = charmm_r_.find(k);
    return it == charmm_r_.end() ? 1.8f : it->second;
}
float Parameter::atom_charge(const std::string &k) const {
    auto it = charmm_c_.find(k);
    return it == charmm_c_.end() ? 0.0f : it->second;
}
float Parameter::atom_ace(const std::string &k) const {
    auto it = ace_.find(k);
    return it == ace_.end() ? 0.0f : it->second;
}

// ------------------------------------------------------------------ PDB
std::string PDBEntry::erase_space(const std::string &s0) {
    std::string s = s0;
    for (size_t n; (n = s.find("  ")) != std::string::npos;) s.erase(n, 1);
    if (!s.empty() && s[0] == ' ') s.erase(0, 1);
    if (!s.empty() && s[s.size() - 1] == ' ') s.erase(s.size() - 1, 1);
    return s;
}

int PDBEntry::pdb_read(std::string &err) {
    std::ifstream in(file_);
    if (!in) { err = "[ERROR] PDB file [" + file_ + "] not open!!"; return 1; }
    std::string line;
    bool warned = false;
    auto field = [&](size_t pos, size_t len) { return pos < line.size() ? line.substr(pos, len) : std::string(); };
    while (std::getline(in, line)) {
        const std::string rec = field(0, 6);
        if (rec != "ATOM  " && rec != "HETATM") continue;
        if (field(76, 2) == " H") {
            if (!warned) std::printf("# This PDB file contains hydrogen atom. MEGADOCK skips hydrogen atom line.\n");
            warned = true;
            continue;
        }
        for (size_t c = 30; c <= 46; c += 8) xyz_.push_back((float)std::atof(field(c, 8).c_str()));
        keys_.push_back(erase_space(field(11, 10)));
    }
    return 0;
}

// ------------------------------------------------------------------ Protein
int Protein::initialize(const Parameter &p, std::string &err) {
    if (pdb_read(err)) return 1;
    const int na = num_atoms();
    if (na < 1) { err = "no atoms in " + file_; return 1; }
    for (int j = 0; j < 3; j++) {  // src/protein.cpp:67-111
        edge_[j][0] = 1.0e9f;
        edge_[j][1] = -1.0e9f;
    }
    for (int i = 0; i < na; i++)
        for (int j = 0; j < 3; j++) {
            const float c = coordinate(i, j);
            if (c < edge_[j][0]) edge_[j][0] = c;
            if (c > edge_[j][1]) edge_[j][1] = c;
        }
    for (int j = 0; j < 3; j++) center_[j] = (float)((double)(edge_[j][0] + edge_[j][1]) / 2.0);
    radius.resize(na);
    charge.resize(na);
    ace.resize(na);
    for (int i = 0; i < na; i++) {
        radius[i] = p.atom_radius(atom_type(i));
        charge[i] = p.atom_charge(atom_type(i));
        ace[i] = p.atom_ace(atom_type(i));
    }
    return 0;
}

void Protein::shift_center(float *out) const {
    for (int i = 0; i < num_atoms(); i++)
        for (int j = 0; j < 3; j++) out[3 * i + j] = coordinate(i, j) - center_[j];
}

std::vector<float> grid_coord(int ng, float gw) {
    std::vector<float> g(ng);
    const float search_length = gw * ng;
    for (int i = 0; i < ng; i++) g[i] = (float)(search_length * (-0.5 + (0.5 + i) / ng));
    return g;
}

// ------------------------------------------------------------------ grid sizing
static const std::vector<int> &grid_table(int set) {  // src/control.cpp:91-285
    static const std::vector<int> t11 = {
        2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 14, 15, 16, 18, 20, 21, 22, 24, 25, 27, 28, 30, 32, 33, 35, 36,
        40, 42, 44, 45, 48, 49, 50, 54, 55, 56, 60, 63, 64, 66, 70, 72, 75, 77, 80, 81, 84, 88, 90, 96, 98,
        99, 100, 105, 108, 110, 112, 120, 121, 125, 126, 128, 132, 135, 140, 144, 147, 150, 154, 160, 162,
        165, 168, 175, 176, 180, 189, 192, 196, 198, 200, 210, 216, 220, 224, 225, 231, 240, 242, 243, 245,
        250, 252, 256, 264, 270, 275, 280, 288, 294, 297, 300, 308, 315, 320, 324, 330, 336, 343, 350, 352,
        360, 363, 375, 378, 384, 385, 392, 396, 400, 405, 420, 432, 440, 441, 448, 450, 462, 480, 484, 486,
        490, 495, 500, 504, 512, 525, 528, 539, 540, 550, 560, 567, 576, 588, 594, 600, 605, 616, 625, 630,
        640, 648, 660, 672, 675, 686, 693, 700, 704, 720, 726, 729, 735, 750, 756, 768, 770, 784, 792, 800,
        810, 825, 840, 847, 864, 875, 880, 882, 891, 896, 900, 924, 945, 960, 968, 972, 980, 990, 1000, 1008,
        1024};
    static const std::vector<int> t13 = {
        2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 18, 20, 21, 22, 24, 25, 26, 27, 28, 30, 32, 33,
        35, 36, 39, 40, 42, 44, 45, 48, 49, 50, 52, 54, 55, 56, 60, 63, 64, 65, 66, 70, 72, 75, 77, 78, 80,
        81, 84, 88, 90, 91, 96, 98, 99, 100, 104, 105, 108, 110, 112, 117, 120, 121, 125, 126, 128, 130, 132,
        135, 140, 143, 144, 147, 150, 154, 156, 160, 162, 165, 168, 169, 175, 176, 180, 182, 189, 192, 195,
        196, 198, 200, 208, 210, 216, 220, 224, 225, 231, 234, 240, 242, 243, 245, 250, 252, 256, 260, 264,
        270, 273, 275, 280, 286, 288, 294, 297, 300, 308, 312, 315, 320, 324, 325, 330, 336, 338, 343, 350,
        351, 352, 360, 363, 364, 375, 378, 384, 385, 390, 392, 396, 400, 405, 416, 420, 429, 432, 440, 441,
        448, 450, 455, 462, 468, 480, 484, 486, 490, 495, 500, 504, 507, 512, 520, 525, 528, 539, 540, 546,
        550, 560, 567, 572, 576, 585, 588, 594, 600, 605, 616, 624, 625, 630, 637, 640, 648, 650, 660, 672,
        675, 676, 686, 693, 700, 702, 704, 715, 720, 726, 728, 729, 735, 750, 756, 768, 770, 780, 784, 792,
        800, 810, 819, 825, 832, 840, 845, 847, 858, 864, 875, 880, 882, 891, 896, 900, 910, 924, 936, 945,
        960, 968, 972, 975, 980, 990, 1000, 1001, 1008, 1014, 1024};
    static const std::vector<int> t07 = {
        2, 3, 4, 5, 6, 7, 8, 9, 10, 12, 14, 15, 16, 18, 20, 21, 24, 25, 27, 28, 30, 32, 35, 36, 40, 42, 45,
        48, 49, 50, 54, 56, 60, 63, 64, 70, 72, 75, 80, 81, 84, 90, 96, 98, 100, 105, 108, 112, 120, 125, 126,
        128, 135, 140, 144, 147, 150, 160, 162, 168, 175, 180, 189, 192, 196, 200, 210, 216, 224, 225, 240,
        243, 245, 250, 252, 256, 270, 280, 288, 294, 300, 315, 320, 324, 336, 343, 350, 360, 375, 378, 384,
        392, 400, 405, 420, 432, 441, 448, 450, 480, 486, 490, 500, 504, 512, 525, 540, 560, 567, 576, 588,
        600, 625, 630, 640, 648, 672, 675, 686, 700, 720, 729, 735, 750, 756, 768, 784, 800, 810, 840, 864,
        875, 882, 896, 900, 945, 960, 972, 980, 1000, 1008, 1024};
    static const std::vector<int> t00 = {
        8, 10, 16, 18, 21, 22, 24, 25, 32, 33, 35, 36, 39, 40, 42, 44, 45, 49, 50, 52, 55, 56, 60, 64, 70, 75,
        78, 84, 91, 100, 105, 110, 117, 126, 130, 140, 147, 150, 154, 156, 162, 165, 169, 175, 189, 200, 210,
        216, 220, 225, 231, 234, 240, 242, 243, 245, 250, 252, 256, 260, 264, 270, 273, 275, 280, 286, 288,
        294, 297, 300, 308, 312, 315, 320, 324, 325, 330, 336, 338, 343, 350, 351, 352, 360, 363, 364, 375,
        378, 384, 385, 390, 392, 396, 400, 405, 416, 420, 429, 432, 440, 441, 448, 450, 455, 462, 468, 480,
        484, 486, 490, 495, 500, 504, 507, 512, 520, 525, 528, 539, 540, 546, 550, 560, 567, 572, 576, 585,
        588, 594, 600, 605, 616, 624, 625, 630, 637, 640, 648, 650, 660, 672, 675, 676, 686, 693, 700, 702,
        704, 715, 720, 726, 728, 729, 735, 750, 756, 768, 770, 780, 784, 792, 800, 810, 819, 825, 832, 840,
        845, 847, 858, 864, 875, 880, 882, 891, 896, 900, 910, 924, 936, 945, 960, 968, 972, 975, 980, 990,
        1000, 1001, 1008, 1014, 1024};
    static const std::vector<int> t01 = {
        16, 32, 36, 40, 42, 48, 50, 64, 72, 80, 81, 84, 96, 98, 100, 108, 112, 128, 144, 160, 162, 168, 200,
        216, 224, 240, 243, 245, 250, 252, 256, 270, 280, 288, 294, 300, 315, 320, 324, 336, 343, 350, 360,
        375, 378, 384, 392, 400, 405, 420, 432, 441, 448, 450, 480, 486, 490, 500, 504, 512, 525, 540, 560,
        567, 576, 588, 600, 625, 630, 640, 648, 672, 675, 686, 700, 720, 729, 735, 750, 756, 768, 784, 800,
        810, 840, 864, 875, 882, 896, 900, 945, 960, 972, 980, 1000, 1008, 1024};
    switch (set) {
        case 13: return t13;
        case 7: return t07;
        case 0: return t00;
        case 1: return t01;
        default: return t11;
    }
}

static int table_pick(int need, int set) {
    for (int v : grid_table(set))
        if (v >= need) return v;
    return need;
}

static float receptor_extent(const Protein &r) {
    float size_rec = 0.0f;
    for (int i = 0; i < 3; i++) {
        const float s = r.edge(i, 1) - r.edge(i, 0);
        if (s > size_rec) size_rec = s;
    }
    return size_rec;
}

static float ligand_extent(const Protein &l) {  // max pair distance (src/control.cpp:333-352)
    float best = 0.0f;
    const int na = l.num_atoms();
    for (int i = 0; i < na - 1; i++) {
        const float x1 = l.coordinate(i, 0), y1 = l.coordinate(i, 1), z1 = l.coordinate(i, 2);
        for (int j = i + 1; j < na; j++) {
            const float dx = l.coordinate(j, 0) - x1, dy = l.coordinate(j, 1) - y1, dz = l.coordinate(j, 2) - z1;
            const float d2 = dx * dx + dy * dy + dz * dz;
            if (d2 > best) best = d2;
        }
    }
    return std::sqrt(best);
}

int autogrid_receptor(const Protein &r, const Parameter &p) {
    float size = receptor_extent(r);
    size += 2.0 * p.grid_space_rec;
    return table_pick(1 + (int)(size / p.grid_width), p.fft_base_set);
}

int autogrid_ligand(const Protein &l, const Parameter &p) {
    float size = ligand_extent(l);
    size += 2.0 * p.grid_space_lig;
    return table_pick(1 + (int)(size / p.grid_width), p.fft_base_set);
}

int check_grid(const Protein &r, const Protein &l, const Parameter &p, std::string &err) {
    const float search_length = p.grid_width * p.num_grid;
    float sr = receptor_extent(r);
    sr += 2.0 * p.grid_space_rec;
    if (sr > search_length) { err = "[ERROR] Receptor data is too big!!"; return 1; }
    float sl = ligand_extent(l);
    sl += 2.0 * p.grid_space_lig;
    if (sl > search_length) { err = "[ERROR] Ligand data is too big!!"; return 1; }
    return 0;
}

}  // namespace mdk
