// megadock-hip-decoygen: the reference's decoygen tool (src/decoygen.cpp) with the
// coordinates computed on the device (mdk_decoy_coords).  Same command line, same input
// reading (the .out header and decoy line read with fscanf, the ligand PDB read in the
// tool's fgets chunks: 30 characters, three 8-character coordinates, the rest of the line)
// and the same output lines.  `decoy no.` may also be a range first-last, which writes one
// file per decoy, <decoy_filename>.<n>.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../../include/mdk_hip.h"

namespace {

struct OutHeader {  // src/decoygen.cpp:108-122
    int N = 0;
    float spacing = 0, rand[3] = {0, 0, 0}, r[3] = {0, 0, 0}, l[3] = {0, 0, 0};
};

struct Decoy {
    float a[3], t[3], score;
};

int read_out(const char *path, int last, OutHeader &h, std::vector<Decoy> &d) {
    FILE *fp = std::fopen(path, "r");
    if (!fp) return 1;
    char rec[300], lig[300];
    if (std::fscanf(fp, "%d %f", &h.N, &h.spacing) != 2 || std::fscanf(fp, "%f %f %f", &h.rand[0], &h.rand[1], &h.rand[2]) != 3 ||
        std::fscanf(fp, "%299s %f %f %f", rec, &h.r[0], &h.r[1], &h.r[2]) != 4 ||
        std::fscanf(fp, "%299s %f %f %f", lig, &h.l[0], &h.l[1], &h.l[2]) != 4) {
        std::fclose(fp);
        return 1;
    }
    d.resize(last);
    for (int i = 0; i < last; i++) {
        Decoy &x = d[i];
        if (std::fscanf(fp, "%f %f %f %f %f %f %f", &x.a[0], &x.a[1], &x.a[2], &x.t[0], &x.t[1], &x.t[2], &x.score) != 7) {
            std::fclose(fp);
            return 1;
        }
    }
    std::fclose(fp);
    return 0;
}

struct PdbPiece {  // one output piece: an atom (prefix, coordinates, rest) or a TER/END line
    std::string head, tail;
    bool atom;
};

// src/decoygen.cpp:74-101: the ligand PDB in the tool's fgets chunks
int read_ligand(const char *path, const float *l, std::vector<PdbPiece> &pieces, std::vector<float> &xyz) {
    FILE *fp = std::fopen(path, "r");
    if (!fp) return 1;
    char liner[85], c1[10], c2[10], c3[10], last[100];
    while (std::fgets(liner, 31, fp) != nullptr) {
        if (!std::strncmp(liner, "ATOM  ", 6) || !std::strncmp(liner, "HETATM", 6)) {
            if (!std::fgets(c1, 9, fp) || !std::fgets(c2, 9, fp) || !std::fgets(c3, 9, fp)) break;
            if (!std::fgets(last, 60, fp)) last[0] = 0;
            xyz.push_back((float)(std::atof(c1) - l[0]));
            xyz.push_back((float)(std::atof(c2) - l[1]));
            xyz.push_back((float)(std::atof(c3) - l[2]));
            pieces.push_back({liner, last, true});
        } else if (!std::strncmp(liner, "TER   ", 6) || !std::strncmp(liner, "ENDMDL", 6) ||
                   !std::strncmp(liner, "END   ", 6)) {
            if (!std::fgets(last, 60, fp)) last[0] = 0;
            pieces.push_back({liner, last, false});
        }
    }
    std::fclose(fp);
    return 0;
}

}  // namespace

int main(int argc, char **argv) {
    if (argc != 5) {
        std::printf("++ decoygen tool (MI355X) ++\n");
        std::printf("\n%s [decoy_filename] [used_ligand.pdb] [.outfile] [decoy no. | first-last]", argv[0]);
        std::printf("\n  **Note: files must be in PDB format.\n");
        return 1;
    }
    int first = (int)std::atof(argv[4]), last = first;
    const char *dash = std::strchr(argv[4], '-');
    const bool range = dash && dash != argv[4];
    if (range) last = std::atoi(dash + 1);
    if (first < 1 || last < first) {
        std::fprintf(stderr, "decoy numbers start at 1\n");
        return 1;
    }
    OutHeader h;
    std::vector<Decoy> dec;
    if (read_out(argv[3], last, h, dec)) {
        std::fprintf(stderr, "cannot read %d decoys from %s\n", last, argv[3]);
        return 1;
    }
    std::vector<PdbPiece> pieces;
    std::vector<float> xyz;
    if (read_ligand(argv[2], h.l, pieces, xyz)) {
        std::fprintf(stderr, "cannot open %s\n", argv[2]);
        return 1;
    }
    const int na = (int)(xyz.size() / 3), nd = last - first + 1;
    float pre[9];
    std::vector<float> ang(3 * (size_t)nd), rot(9 * (size_t)nd), out(3 * (size_t)na * nd);
    std::vector<int> trans(3 * (size_t)nd);
    for (int q = 0; q < nd; q++) {
        const Decoy &x = dec[first - 1 + q];
        for (int k = 0; k < 3; k++) {
            ang[3 * q + k] = x.a[k];
            int t = (int)x.t[k];  // src/decoygen.cpp:92-94
            if (t >= h.N / 2) t -= h.N;
            trans[3 * q + k] = t;
        }
    }
    if (mdk_rotation_coef(h.rand, 1, pre) || mdk_rotation_coef(ang.data(), nd, rot.data()) ||
        (na > 0 && mdk_decoy_coords(0, xyz.data(), na, pre, rot.data(), trans.data(), nd, h.spacing, h.r, out.data()))) {
        std::fprintf(stderr, "%s\n", mdk_last_error());
        return 1;
    }
    for (int q = 0; q < nd; q++) {
        std::string name = argv[1];
        if (range) name += "." + std::to_string(first + q);
        FILE *fp = std::fopen(name.c_str(), "w");
        if (!fp) {
            std::fprintf(stderr, "cannot write %s\n", name.c_str());
            return 1;
        }
        size_t a = 0;
        for (const PdbPiece &p : pieces) {
            if (p.atom) {
                const float *o = &out[3 * ((size_t)q * na + a++)];
                std::fprintf(fp, "%s%8.3f%8.3f%8.3f%s", p.head.c_str(), o[0], o[1], o[2], p.tail.c_str());
            } else {
                std::fprintf(fp, "%s%s", p.head.c_str(), p.tail.c_str());
            }
        }
        std::fclose(fp);
    }
    return 0;
}
