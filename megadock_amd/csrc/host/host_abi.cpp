// C ABI of the host layer (include/mdk_host.h).
#include <cstring>
#include <string>

#include "../../../include/mdk_host.h"
#include "mdk_host.h"

struct mdkh_protein {
    mdk::Receptor prot;  // a Receptor also serves as a plain Protein (ligand)
    explicit mdkh_protein(const std::string &f) : prot(f) {}
};

namespace {
thread_local std::string g_herr;

int herr(const std::string &m) {
    g_herr = m;
    return 1;
}

const mdk::Parameter *tables(std::string &err) {
    static mdk::Parameter *p = nullptr;
    static std::string load_err;
    static bool tried = false;
    if (!tried) {
        tried = true;
        mdk::Parameter *q = new mdk::Parameter();
        if (q->load_tables(load_err)) delete q;
        else p = q;
    }
    if (!p) err = load_err;
    return p;
}
}  // namespace

extern "C" {

const char *mdkh_last_error(void) { return g_herr.c_str(); }

int mdkh_protein_load(const char *path, mdkh_protein **out) {
    if (!path || !out) return herr("mdkh_protein_load: null argument");
    *out = nullptr;
    std::string err;
    const mdk::Parameter *p = tables(err);
    if (!p) return herr(err);
    mdkh_protein *m = new mdkh_protein(path);
    if (m->prot.initialize(*p, err)) {
        delete m;
        return herr(err);
    }
    *out = m;
    return 0;
}

int mdkh_protein_free(mdkh_protein *p) {
    delete p;
    return 0;
}

int mdkh_protein_size(const mdkh_protein *p, int *na) {
    if (!p || !na) return herr("mdkh_protein_size: null argument");
    *na = p->prot.num_atoms();
    return 0;
}

int mdkh_protein_atoms(const mdkh_protein *p, float *xyz, float *radius, float *charge, float *ace, float *center) {
    if (!p) return herr("mdkh_protein_atoms: null protein");
    const auto &q = p->prot;
    const int na = q.num_atoms();
    if (xyz) std::memcpy(xyz, q.coords().data(), sizeof(float) * 3 * na);
    if (radius) std::memcpy(radius, q.radius.data(), sizeof(float) * na);
    if (charge) std::memcpy(charge, q.charge.data(), sizeof(float) * na);
    if (ace) std::memcpy(ace, q.ace.data(), sizeof(float) * na);
    if (center)
        for (int j = 0; j < 3; j++) center[j] = q.center(j);
    return 0;
}

int mdkh_autogrid(const mdkh_protein *rec, const mdkh_protein *lig, float gw, int table, int *ng) {
    if (!rec || !lig || !ng) return herr("mdkh_autogrid: null argument");
    mdk::Parameter p;
    p.grid_width = gw;
    p.fft_base_set = table;
    const int a = mdk::autogrid_receptor(rec->prot, p);
    const int b = mdk::autogrid_ligand(lig->prot, p);
    *ng = a > b ? a : b;
    return 0;
}

int mdkh_rotation_set(int n, float *angles, int *count) {
    if (!count) return herr("mdkh_rotation_set: null count");
    std::vector<float> a;
    std::string err;
    if (mdk::Parameter::load_rotation_set(n, a, err)) return herr(err);
    *count = (int)(a.size() / 3);
    if (angles) std::memcpy(angles, a.data(), sizeof(float) * a.size());
    return 0;
}

int mdkh_grid_coord(int ng, float gw, float *out) {
    if (ng < 1 || !out) return herr("mdkh_grid_coord: invalid argument");
    const auto g = mdk::grid_coord(ng, gw);
    std::memcpy(out, g.data(), sizeof(float) * ng);
    return 0;
}

int mdkh_receptor_voxel(const mdkh_protein *rec, int ng, float gw, float aceratio, float rec_core, float eratio,
                        float *grid_r, float *grid_i) {
    if (!rec || ng < 2 || !grid_r || !grid_i) return herr("mdkh_receptor_voxel: invalid argument");
    const auto &r = rec->prot;
    std::vector<float> xyz(3 * (size_t)r.num_atoms());
    r.shift_center(xyz.data());
    r.voxelize(ng, gw, mdk::grid_coord(ng, gw), xyz.data(), aceratio, rec_core, eratio, grid_r, grid_i);
    return 0;
}

int mdkh_run(int argc, const char **argv) {
    std::string err;
    const int rc = mdk::run(argc, const_cast<char **>(argv), err);
    if (rc) return herr(err);
    return 0;
}

int mdkh_run_table(int argc, const char **argv) {
    std::string err;
    const int rc = mdk::run_table(argc, const_cast<char **>(argv), err);
    if (rc) return herr(err);
    return 0;
}

}  // extern "C"
