// Driver around the reference's own CPU sources -- TEST INFRASTRUCTURE ONLY.
//
// Compiled together with the reference's src/{parameter,pdb_entry,protein,receptor,
// ligand,decoygen}.cpp (read where they lie, never copied) into oracle/_ref/libmdkref.so.
// It exposes a C ABI so the tests can dump golden grids from the reference's own
// Receptor/Ligand voxelizers, its PDB reader and parameter tables, and its ZXZ rotation
// (decoygen.cpp's rotateAtom, the same float expression as src/docking.cpp:376-393).
// Nothing in the product links this file.
#include <getopt.h>

#include <cstring>
#include <string>
#include <vector>

#include "parameter.h"
#include "receptor.h"
#include "ligand.h"

void rotateAtom(float, float, float, float *, float *, float *, float, float, float);

// src/parameter.h:39 names HeaderDB a friend of Parameter and the reference defines it
// nowhere; defining it here gives the tests the rotation table exactly as
// Parameter::parameter_set loads it (_Zangle, _Num_rot_angles: src/parameter.h:87-88).
class HeaderDB {
public:
    static const float *angles(const Parameter *p) { return p->_Zangle; }
    static int count(const Parameter *p) { return p->_Num_rot_angles; }
};

namespace {
Parallel *g_par = nullptr;
Parameter *g_prm = nullptr;
Receptor *g_rec = nullptr;
Ligand *g_lig = nullptr;
std::string g_rec_file, g_lig_file;

Protein *pick(int which) { return which == 0 ? static_cast<Protein *>(g_rec) : g_lig; }

// grid coordinates as src/docking.cpp:133-141 computes them
void grid_coord(int ng, float gw, float *gc) {
    float search_length = gw * ng;
    for (int i = 0; i < ng; i++) gc[i] = search_length * (-0.5 + (0.5 + i) / ng);
}
}  // namespace

extern "C" {

// argv-style options after -R/-L, e.g. {"-v","1.2"}; returns 0 on success.
int ref_load(const char *rec_pdb, const char *lig_pdb, int nextra, const char **extra) {
    g_rec_file = rec_pdb;
    g_lig_file = lig_pdb;
    std::vector<std::string> args = {"megadock", "-R", rec_pdb, "-L", lig_pdb};
    for (int i = 0; i < nextra; i++) args.push_back(extra[i]);
    std::vector<char *> argv;
    for (auto &s : args) argv.push_back(const_cast<char *>(s.c_str()));
    delete g_rec; delete g_lig; delete g_prm; delete g_par;
    g_par = new Parallel(1);
    g_prm = new Parameter(g_par);
    optind = 0;  // getopt restarts from argv[1] on every load (glibc: 0 reinitialises)
    g_prm->initialize((int)argv.size(), argv.data());
    g_rec = new Receptor(g_rec_file);
    g_rec->initialize(g_prm);
    g_lig = new Ligand(g_lig_file);
    g_lig->initialize(g_prm);
    return 0;
}

int ref_num_atoms(int which) { return pick(which)->num_atoms(); }

void ref_coords(int which, float *out) {
    Protein *p = pick(which);
    for (int i = 0; i < p->num_atoms(); i++)
        for (int j = 0; j < 3; j++) out[3 * i + j] = p->coordinate(i, j);
}

// radius, charge, ACE per atom through the reference's table lookups
void ref_atom_params(int which, float *radius, float *charge, float *ace) {
    Protein *p = pick(which);
    for (int i = 0; i < p->num_atoms(); i++) {
        radius[i] = g_prm->atom_radius(p->atom_type(i));
        charge[i] = g_prm->atom_charge(p->atom_type(i));
        ace[i] = g_prm->atom_ace(p->atom_type(i));
    }
}

// atom-type key of atom i into buf (NUL-terminated, at most n bytes)
void ref_atom_type(int which, int i, char *buf, int n) {
    std::string s = pick(which)->atom_type(i);
    std::strncpy(buf, s.c_str(), n - 1);
    buf[n - 1] = 0;
}

void ref_center(int which, float *c) {
    for (int j = 0; j < 3; j++) c[j] = pick(which)->center(j);
}

void ref_edge(int which, float *e) {
    for (int j = 0; j < 3; j++) { e[2 * j] = pick(which)->edge(j, 0); e[2 * j + 1] = pick(which)->edge(j, 1); }
}

static void work_buffers(int ng, int na, std::vector<float> &fw, std::vector<int> &iw) {
    const size_t ng3 = (size_t)ng * ng * ng, nag = (size_t)na * ng;
    fw.assign(ng3 * 3 + (size_t)na * 3 + nag * 3, 0.0f);
    iw.assign(ng3 * 2 + (size_t)na * 4, 0);
}

// Receptor grids as src/docking.cpp:128-191 builds them (shift_center + rpscace + electro)
void ref_receptor_voxel(int ng, float gw, float aceratio, float rec_core, float lig_core,
                        float eratio, float *grid_r, float *grid_i) {
    const int na = g_rec->num_atoms();
    const size_t ng3 = (size_t)ng * ng * ng;
    std::vector<float> gc(ng), mol(3 * (size_t)na), fw;
    std::vector<int> iw;
    grid_coord(ng, gw, gc.data());
    g_rec->shift_center(mol.data());
    work_buffers(ng, na, fw, iw);
    g_rec->rpscace(aceratio, ng, gc.data(), mol.data(), iw.data(), fw.data(), rec_core, lig_core, 0);
    g_rec->electro(-2800.0f, eratio, ng, gc.data(), mol.data(), iw.data(), fw.data(), 0);
    std::memcpy(grid_r, fw.data(), sizeof(float) * ng3);
    std::memcpy(grid_i, fw.data() + ng3, sizeof(float) * ng3);
}

// Ligand grids for already-rotated coordinates (src/ligand.cpp:20-239)
void ref_ligand_voxel(int ng, float gw, const float *xyz_rot, float aceratio, float rec_core,
                      float lig_core, float eratio, float *grid_r, float *grid_i) {
    const int na = g_lig->num_atoms();
    const size_t ng3 = (size_t)ng * ng * ng;
    std::vector<float> gc(ng), mol(xyz_rot, xyz_rot + 3 * (size_t)na), fw;
    std::vector<int> iw;
    grid_coord(ng, gw, gc.data());
    work_buffers(ng, na, fw, iw);
    g_lig->rpscace(aceratio, ng, gc.data(), mol.data(), iw.data(), fw.data(), rec_core, lig_core, 0);
    g_lig->electro(-2800.0f, eratio, ng, gc.data(), mol.data(), iw.data(), fw.data(), 0);
    std::memcpy(grid_r, fw.data(), sizeof(float) * ng3);
    std::memcpy(grid_i, fw.data() + ng3, sizeof(float) * ng3);
}

// the loaded rotation set (-r / -D): number of rotations, and 3 floats each into out
int ref_rotation_count(void) { return HeaderDB::count(g_prm); }

void ref_rotation_angles(float *out) {
    const float *a = HeaderDB::angles(g_prm);
    std::memcpy(out, a, sizeof(float) * 3 * (size_t)HeaderDB::count(g_prm));
}

// decoygen's rotateAtom: (x,y,z) rotated by ZXZ angles (a0, a1, a2)
void ref_rotate(const float *xyz, int n, const float *theta, float *out) {
    for (int i = 0; i < n; i++)
        rotateAtom(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], &out[3 * i], &out[3 * i + 1],
                   &out[3 * i + 2], theta[0], theta[1], theta[2]);
}
}
