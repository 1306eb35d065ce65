// Test-infrastructure driver for the real SPILADY CPU reference.
//
// This file is compiled together with the reference's own sources where they
// lie (no copies) into oracle/_ref/libspilady_ref.so by oracle/Makefile.ref.
// It is a checker: only tests, the smoke check and bench.py's cpu_baseline leg
// load it. It never ships in the product path.
//
// The reference keeps all state in globals (global.h:53-214) and is driven by
// main() in spilady.cpp:33-72. That main() is renamed at compile time so this
// driver can call the same functions in the same order with a fixed step
// count and without scale() (scale_step, the adaptive dt, lives only there:
// scale.cpp:41-43).
//
// Two ways in:
//   ref_init_files  - the reference's own initialize() (initial.cpp:31-106):
//                     variables.in + con- file read from a directory;
//   ref_init_memory - the same sequence, but the atoms come from arrays
//                     instead of read_config() (read_config.cpp:32-147), for
//                     sizes where a text con- file is impractical.

#include "spilady.h"

#include <streambuf>
#include <unistd.h>
#include <string>
#include <vector>

namespace {

struct NullBuf : public std::streambuf {
    int overflow(int c) override { return c; }
};
NullBuf g_null;
std::streambuf* g_cout_saved = nullptr;
std::string g_dir;
bool g_live = false;

struct Quiet {
    char old[4096];
    bool moved = false;
    explicit Quiet(bool to_dir) {
        g_cout_saved = std::cout.rdbuf(&g_null);
        if (to_dir && !g_dir.empty()) {
            if (getcwd(old, sizeof(old)) && chdir(g_dir.c_str()) == 0) moved = true;
        }
    }
    ~Quiet() {
        if (moved) { if (chdir(old) != 0) {} }
        std::cout.rdbuf(g_cout_saved);
    }
};

inline int aid(const atom_struct* a) { return int(a - first_atom_ptr); }

}  // namespace

extern "C" {

// Reference initialize() with readconf (initial.cpp:31-106), then check(-1),
// as spilady.cpp:38-39 does. The directory must hold variables.in and the
// con- file it names. Output files (enr-, tmp-, spn-) are written there.
int ref_init_files(const char* dir) {
    g_dir = dir;
    Quiet q(true);
    initialize();
    check(-1);
    g_live = true;
    return natom;
}

// Same order as initialize() (initial.cpp:31-106), with read_config()
// (read_config.cpp:32-147) replaced by arrays. variables.in is still parsed
// by the reference's read_variables() from `dir`.
// Arrays are [natom][3] for vectors; stress is [natom][6] in the con- column
// order 11 22 33 12 23 31.
int ref_init_memory(const char* dir, int n, const double* dbox, const double* r,
                    const double* p, const double* f, const double* stress,
                    const double* rho, const double* ke, const double* pe,
                    const double* s, const double* me, const double* me0,
                    double start_time_in) {
    g_dir = dir;
    Quiet q(true);
    read_variables();
    omp_set_num_threads(OMP_threads);
    potential_table();

    natom = n;
    total_time = start_time_in;
    d.xx = dbox[0]; d.yx = dbox[1]; d.yy = dbox[2];
    d.zx = dbox[3]; d.zy = dbox[4]; d.zz = dbox[5];
    Inv_d = inverse_box_vector(d);
    first_atom_ptr = (atom_struct*)calloc(natom, sizeof(atom_struct));
    for (int i = 0; i < natom; ++i) {
        atom_struct* a = first_atom_ptr + i;
        a->element[0] = 'F'; a->element[1] = 'e';
        a->r.x = r[3*i]; a->r.y = r[3*i+1]; a->r.z = r[3*i+2];
        a->p.x = p[3*i]; a->p.y = p[3*i+1]; a->p.z = p[3*i+2];
        a->f.x = f[3*i]; a->f.y = f[3*i+1]; a->f.z = f[3*i+2];
        a->stress11 = stress[6*i];   a->stress22 = stress[6*i+1];
        a->stress33 = stress[6*i+2]; a->stress12 = stress[6*i+3];
        a->stress23 = stress[6*i+4]; a->stress31 = stress[6*i+5];
        a->rho = rho[i]; a->ke = ke[i]; a->pe = pe[i];
        a->s.x = s[3*i]; a->s.y = s[3*i+1]; a->s.z = s[3*i+2];
        a->s0 = vec_length(a->s);
        a->me = me[i]; a->me0 = me0[i];
        a->local_volume = 0e0;
    }
    box_length.x = fabs(d.xx);
    box_length.y = sqrt(d.yx*d.yx + d.yy*d.yy);
    box_length.z = sqrt(d.zx*d.zx + d.zy*d.zy + d.zz*d.zz);
    box_length_half = vec_times(0.5, box_length);
    box_volume = d.xx*d.yy*d.zz;
    density = double(natom)/box_volume;
    // read_config.cpp:133-145: with magmom the input column holds the moments
    for (int i = 0; i < natom; ++i) {
        atom_struct* a = first_atom_ptr + i;
#ifdef magmom
        a->m = a->s;
        a->m0 = vec_length(a->m);
        a->s = vec_divide(a->m, -el_g);
        a->s0 = vec_length(a->s);
#else
        a->m = vec_times(-el_g, a->s);
        a->m0 = vec_length(a->m);
#endif
    }
    // initial.cpp:58-63: extfield.in / extforce.in of the run directory
#ifdef extfield
    external_field(-1);
#endif
#ifdef extforce
    external_force(-1);
#endif
    map_cells();
    initial_links();
    allocate_cells();
    calculate_temperature();
    total_time = start_time;
    // initial.cpp:93-98: the barostats' clocks start at the start time
#ifdef PRESSURE
    last_total_time_pressure = total_time;
#endif
#ifdef STRESS
    last_total_time_stress = total_time;
#endif
    initialize_random_number();
    // initial.cpp:102-104: the quantum thermostat's filter and noise history
#ifdef quantumnoise
    initial_quantum_noise();
#endif
    check(-1);
    g_live = true;
    return natom;
}

// One reference time step as the driver loop does it (spilady.cpp:50-57),
// without scale() and write().
void ref_run(int first_step, int nsteps, int do_check) {
    Quiet q(true);
    for (int k = 0; k < nsteps; ++k) {
        int cur = first_step + k;
        core(cur);
        total_time += step;
        if (do_check) check(cur);
    }
}

// Single reference entry points (prototype_CPU.h:120-289), for per-kernel
// fixtures.
void ref_links(void) { Quiet q(false); links(); }
void ref_calculate_rho(void) { Quiet q(false); calculate_rho(); }
void ref_calculate_force_energy(void) { Quiet q(false); calculate_force_energy(); }
void ref_core_ds(double dt) { Quiet q(false); core_ds(dt); }
void ref_core_dp(double dt) { Quiet q(false); core_dp(dt); }
void ref_core_dr(double dt) { Quiet q(false); core_dr(dt); }
void ref_core_dTe(double dt) { Quiet q(false); core_dTe(dt); }
void ref_calculate_temperature(void) { Quiet q(false); calculate_temperature(); }
void ref_check(int cur) { Quiet q(true); check(cur); }
// scale() (scale.cpp:31-44): with the shipped control.h, scale_step only;
// a build with -DSTRESS -DPRESSURE adds the two barostats before it
void ref_scale(int cur) { Quiet q(true); scale(cur); }
// the quantum thermostat's filter H (quantum_noise_CPU.cpp:83-113), 0 without it
int ref_quantum_taps(void) {
#ifdef quantumnoise
    return Nfrequency_quantum_2;
#else
    return 0;
#endif
}
void ref_quantum_H(double* out) {
#ifdef quantumnoise
    for (int i = 0; i < Nfrequency_quantum_2; ++i) out[i] = H_ptr[i];
#else
    (void)out;
#endif
}
// write_vsim(step) (write_vsim.cpp:33-134; the shipped control.h has writevsim)
void ref_write_vsim(int cur) {
#ifdef writevsim
    Quiet q(true);
    write_vsim(cur);
#else
    (void)cur;
#endif
}
void ref_get_box(double* b) { b[0] = d.xx; b[1] = d.yx; b[2] = d.yy; b[3] = d.zx; b[4] = d.zy; b[5] = d.zz; }

int ref_natom(void) { return natom; }
int ref_ncells(void) { return ncells; }
int ref_ngroups(void) { return ngroups; }
void ref_cell_dims(int* n3) {
    n3[0] = no_of_link_cell_x; n3[1] = no_of_link_cell_y; n3[2] = no_of_link_cell_z;
}
double ref_step(void) { return step; }
double ref_total_time(void) { return total_time; }
double ref_virial(void) { return virial; }
double ref_initial_ave_energy(void) { return initial_ave_energy; }
double ref_scalar(const char* name) {
    std::string k(name);
    if (k == "atmass") return atmass;
    if (k == "temperature") return temperature;
    if (k == "gamma_L") return gamma_L;
    if (k == "gamma_L_over_mass") return gamma_L_over_mass;
    if (k == "gamma_S_HL") return gamma_S_HL;
    if (k == "rhomax") return rhomax;
    if (k == "rmax") return rmax;
    if (k == "finput") return finput;
    if (k == "min_length_link_cell") return min_length_link_cell;
    if (k == "kappa_e") return kappa_e;
    if (k == "box_volume") return box_volume;
    return 0.0 / 0.0;
}

// Table sample: interpolated value of table `which` at x, through the
// reference's own lookup (potential_CPU.cpp:172-179).
double ref_table(int which, double x) {
    switch (which) {
        case 0: return bigf(x);
        case 1: return dbigf(x);
        case 2: return smallf(x);
        case 3: return dsmallf(x);
        case 4: return pairij(x);
        case 5: return dpair(x);
        case 6: return Jij(x);
        case 7: return dJij(x);
        case 8: return LandauA(x);
        case 9: return LandauB(x);
        case 10: return LandauC(x);
        case 11: return LandauD(x);
        case 12: return dLandauA(x);
        case 13: return dLandauB(x);
        case 14: return dLandauC(x);
        case 15: return dLandauD(x);
    }
    return 0.0 / 0.0;
}

// Raw table entry n of table `which` (potential_CPU.cpp:106-156).
double ref_table_raw(int which, int n) {
    double* t[16] = {bf_ptr, dbf_ptr, sf_ptr, dsf_ptr, pr_ptr, dpr_ptr, Jij_ptr, dJij_ptr,
                     LandauA_ptr, LandauB_ptr, LandauC_ptr, LandauD_ptr,
                     dLandauA_ptr, dLandauB_ptr, dLandauC_ptr, dLandauD_ptr};
    if (which < 0 || which > 15 || n < 0 || n >= ninput) return 0.0 / 0.0;
    return t[which][n];
}

// Atom state, [natom][3] vectors, stress [natom][6].
void ref_get_atoms(double* r, double* p, double* f, double* s, double* heff_h,
                   double* heff_l, double* rho, double* pe, double* ke, double* me,
                   double* me0, double* stress, double* vir, double* sum_jij_sj,
                   int* cell) {
    for (int i = 0; i < natom; ++i) {
        const atom_struct* a = first_atom_ptr + i;
        r[3*i] = a->r.x; r[3*i+1] = a->r.y; r[3*i+2] = a->r.z;
        p[3*i] = a->p.x; p[3*i+1] = a->p.y; p[3*i+2] = a->p.z;
        f[3*i] = a->f.x; f[3*i+1] = a->f.y; f[3*i+2] = a->f.z;
        s[3*i] = a->s.x; s[3*i+1] = a->s.y; s[3*i+2] = a->s.z;
        heff_h[3*i] = a->Heff_H.x; heff_h[3*i+1] = a->Heff_H.y; heff_h[3*i+2] = a->Heff_H.z;
        heff_l[3*i] = a->Heff_L.x; heff_l[3*i+1] = a->Heff_L.y; heff_l[3*i+2] = a->Heff_L.z;
        rho[i] = a->rho; pe[i] = a->pe; ke[i] = a->ke; me[i] = a->me; me0[i] = a->me0;
        stress[6*i] = a->stress11; stress[6*i+1] = a->stress22; stress[6*i+2] = a->stress33;
        stress[6*i+3] = a->stress12; stress[6*i+4] = a->stress23; stress[6*i+5] = a->stress31;
        vir[i] = a->vir;
        sum_jij_sj[i] = a->sum_Jij_sj;
        cell[i] = a->new_cell_index;
    }
}

// Each atom's local volume (the last column of write_config.cpp:129).
void ref_get_local_volume(double* lv) {
    for (int i = 0; i < natom; ++i) lv[i] = (first_atom_ptr + i)->local_volume;
}

// Each atom's magnetic moment m (struct.h:76), [natom][3].
void ref_get_m(double* m) {
    for (int i = 0; i < natom; ++i) {
        const atom_struct* a = first_atom_ptr + i;
        m[3*i] = a->m.x; m[3*i+1] = a->m.y; m[3*i+2] = a->m.z;
    }
}

// Overwrite the dynamic atom fields (to re-synchronise a run).
void ref_set_atoms(const double* r, const double* p, const double* f, const double* s,
                   const double* rho) {
    for (int i = 0; i < natom; ++i) {
        atom_struct* a = first_atom_ptr + i;
        a->r.x = r[3*i]; a->r.y = r[3*i+1]; a->r.z = r[3*i+2];
        a->p.x = p[3*i]; a->p.y = p[3*i+1]; a->p.z = p[3*i+2];
        a->f.x = f[3*i]; a->f.y = f[3*i+1]; a->f.z = f[3*i+2];
        a->s.x = s[3*i]; a->s.y = s[3*i+1]; a->s.z = s[3*i+2];
#ifdef magmom
        // the spin column holds the moments (read_config.cpp:136-139)
        a->m = a->s;
        a->m0 = vec_length(a->m);
        a->s = vec_divide(a->m, -el_g);
#endif
        a->s0 = vec_length(a->s);
        a->rho = rho[i];
    }
}

// Per-cell linked lists, head to tail, as atom indices. `lists` holds cap
// entries per cell; returns the largest cell population (so the caller can
// detect truncation).
int ref_get_cell_lists(int cap, int* counts, int* lists) {
    int maxn = 0;
    for (int c = 0; c < ncells; ++c) {
        const cell_struct* cp = first_cell_ptr + c;
        int k = 0;
        for (const atom_struct* a = cp->head_ptr; a != NULL; a = a->next_atom_ptr) {
            if (k < cap) lists[c*cap + k] = aid(a);
            ++k;
        }
        counts[c] = k;
        if (k > maxn) maxn = k;
    }
    return maxn;
}

// Group structure from allocate_cells (allocate_cells_CPU.cpp:33-125):
// sizes[g] and the member cells of each group, concatenated in order.
int ref_get_groups(int* sizes, int* members) {
    int k = 0;
    for (int g = 0; g < ngroups; ++g) {
        sizes[g] = allocate_threads_ptr[g];
        for (int j = 0; j < sizes[g]; ++j)
            members[k++] = int(allocate_cell_ptr_ptr[g*max_no_of_members + j] - first_cell_ptr);
    }
    return k;
}

// Per-cell electron subsystem fields (struct.h:122-166).
void ref_get_cells(double* te, double* ee, double* tl, double* ts_r, double* ts_l,
                   double* gel, double* ges, double* sum_r_up, double* sum_r_dn,
                   double* sum_l_up, double* sum_l_dn) {
    for (int c = 0; c < ncells; ++c) {
        const cell_struct* cp = first_cell_ptr + c;
        te[c] = cp->Te; ee[c] = cp->Ee; tl[c] = cp->Tl; ts_r[c] = cp->Ts_R; ts_l[c] = cp->Ts_L;
        gel[c] = cp->Gel; ges[c] = cp->Ges;
        sum_r_up[c] = cp->sum_R_up; sum_r_dn[c] = cp->sum_R_dn;
        sum_l_up[c] = cp->sum_L_up; sum_l_dn[c] = cp->sum_L_dn;
    }
}

void ref_set_cells_te(const double* te) {
    for (int c = 0; c < ncells; ++c) (first_cell_ptr + c)->Te = te[c];
}

void ref_free(void) {
    if (!g_live) return;
    Quiet q(false);
    free_memory();
    first_atom_ptr = NULL;
    first_cell_ptr = NULL;
    g_live = false;
}

}  // extern "C"
