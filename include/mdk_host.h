/*
 * libmdk host layer -- the input side, receptor producer and driver of the drop-in,
 * behind a C ABI (host code only; no GPU is touched except by mdkh_run).
 *
 * All functions return 0 on success and nonzero otherwise; mdkh_last_error() then
 * describes the failure (the reference prints and calls exit()).
 */
#ifndef MDK_HOST_H
#define MDK_HOST_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mdkh_protein mdkh_protein;

/* Read a PDB file and look up per-atom radius, charge and ACE: PDBEntry::pdb_read
 * (src/pdb_entry.cpp:18-75) + Protein::initialize (src/protein.cpp:19-64). */
int mdkh_protein_load(const char *pdb_path, mdkh_protein **out);
int mdkh_protein_free(mdkh_protein *p);
int mdkh_protein_size(const mdkh_protein *p, int *num_atoms);

/* Atom arrays (xyz: na*3, radius/charge/ace: na) and the box centre (3) of
 * Protein::axis_edge (src/protein.cpp:67-111).  Any output pointer may be NULL. */
int mdkh_protein_atoms(const mdkh_protein *p, float *xyz, float *radius, float *charge, float *ace,
                       float *center);

/* Auto grid sizing, Control::autogridr/autogridl (src/control.cpp:287-379) with FFT
 * size table `table` (-l: 11, 13, 7, 0 or 1): *ng = max(receptor ng, ligand ng). */
int mdkh_autogrid(const mdkh_protein *rec, const mdkh_protein *lig, float gw, int table, int *ng);

/* Built-in rotation set n (1, 3, 24, 360, 3600 or 54000, the -r flag): *count = n and,
 * when angles is not NULL, the ZXZ Euler triples (3n floats, rad) as
 * Parameter::parameter_set loads them into _Zangle (src/parameter.cpp:376-395,
 * :3793-61886). */
int mdkh_rotation_set(int n, float *angles, int *count);

/* Grid coordinates, Docking::rec_init (src/docking.cpp:133-141). */
int mdkh_grid_coord(int ng, float gw, float *out);

/* Receptor shape and electrostatics grids (ng^3 each) from the receptor's centred
 * atoms: Protein::shift_center + Receptor::rpscace + Receptor::electro
 * (src/protein.cpp:114-131, src/receptor.cpp:20-372). */
int mdkh_receptor_voxel(const mdkh_protein *rec, int ng, float gw, float aceratio, float rec_core,
                        float eratio, float *grid_r, float *grid_i);

/* The whole program with the reference's command line (src/main.cpp:38-140):
 * reads the PDBs, sizes the grids, docks on GPU 0 and writes the .out file. */
int mdkh_run(int argc, const char **argv);

/* The job-table farm with the reference's table format and options (src/mpidp.cpp:
 * -tb table, -lg log, -rt retries, -ot output field; TITLE= / PARAM= lines, $1..$n
 * replaced by each job line's tab-separated fields), run in one process: host threads
 * prepare jobs (-pj threads) while GPU 0 docks them.  -dry writes the expanded command
 * lines to the log and runs nothing.  Nonzero when any job failed after its retries. */
int mdkh_run_table(int argc, const char **argv);

const char *mdkh_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
