// spilady_hip_bridge.cpp - SPILADY's backend functions on libspilady_hip.so.
//
// SPILADY chooses its backend at link time: each hot-path function declared
// in prototype_CPU.h:120-289 is defined once in a *_CPU.cpp file (built with
// -DCPU) and once in a *_GPU.cu file (-DGPU). This file is a third definition
// of the same functions, compiled with -DHIPLIB, which calls the C ABI of
// include/spilady_hip.h. SPILADY's own main loop (spilady.cpp:33-72),
// initialize() (initial.cpp:31-106), core() (core.cpp:31-76), check(),
// scale() and write() run unchanged and call into it.
//
// Build: every SPILADY .cpp and this file with
//   g++ -O3 -fopenmp -DHIPLIB -DSLDHL -include spilady_hip_bridge.h
//       -I<this repo>/include -I<this repo>/spilady_amd/csrc
// linked with -L<this repo>/spilady_amd -lspilady_hip, after the five lines
// "#ifdef GPU" that call the GPU lifecycle functions (initial.cpp:67,
// write_config.cpp:35 and :137, write_vsim.cpp:37, free_memory.cpp:41) are
// widened to "#if defined GPU || defined HIPLIB". oracle/build_hiplib.sh
// does exactly this, applying the widening as it compiles.
//
// State: the device copy (one sld_ctx) is authoritative during the run, as
// the reference's device copy is in its GPU build (copy_CPU_btw_GPU.cu);
// the host atom and cell arrays are refreshed before output. The clock, the
// time step and the box stay host globals, as in the reference; the bridge
// hands them to the library where the library acts on them.
//
// Errors: the reference's functions return nothing, so a failing call
// prints the library's message and exits, as the reference does on bad input.

#if defined HIPLIB

#include <cstdio>
#include <cstdlib>

// The switches of control.h (spinlang, eltemp, changestep, ...) are macros,
// defined there or on the command line, and those names are also sld_params
// fields: the header and the code that sets the fields see them undefined.
#pragma push_macro("spinlang")
#pragma push_macro("lattlang")
#pragma push_macro("eltemp")
#pragma push_macro("localcolmot")
#pragma push_macro("changestep")
#pragma push_macro("extfield")
#pragma push_macro("extforce")
#pragma push_macro("quantumnoise")
#pragma push_macro("localvol")
#pragma push_macro("magmom")
#undef spinlang
#undef lattlang
#undef eltemp
#undef localcolmot
#undef changestep
#undef extfield
#undef extforce
#undef quantumnoise
#undef localvol
#undef magmom
#include "spilady_hip.h"

namespace {

// control.h switches, in this order, into sld_params; defined here, where
// the field names are not yet macros
enum { kSpinlang, kLattlang, kEltemp, kLocalcolmot, kRenormalize, kChangestep, kExtfield, kExtforce,
       kQuantumnoise, kLocalvol, kMagmom, kNSwitch };
void apply_switches(sld_params& p, const int (&on)[kNSwitch]) {
    p.spinlang = on[kSpinlang];
    p.lattlang = on[kLattlang];
    p.eltemp = on[kEltemp];
    p.localcolmot = on[kLocalcolmot];
    p.renormalize_energy = on[kRenormalize];
    p.changestep = on[kChangestep];
    p.extfield = on[kExtfield];
    p.extforce = on[kExtforce];
    p.quantumnoise = on[kQuantumnoise];
    p.localvol = on[kLocalvol];
    p.magmom = on[kMagmom];
}

}  // namespace

#pragma pop_macro("spinlang")
#pragma pop_macro("lattlang")
#pragma pop_macro("eltemp")
#pragma pop_macro("localcolmot")
#pragma pop_macro("changestep")
#pragma pop_macro("extfield")
#pragma pop_macro("extforce")
#pragma pop_macro("quantumnoise")
#pragma pop_macro("localvol")
#pragma pop_macro("magmom")

#include "spilady.h"

namespace {

sld_ctx* g_ctx = nullptr;
sld_obs g_obs;             // the record of the step being checked
int g_obs_step = -2;       // step of g_obs (-1 is the initial check)

void ok(int rc, const char* what) {
    if (rc != 0) {
        std::fprintf(stderr, "libspilady_hip: %s: %s\n", what, sld_last_error(g_ctx));
        std::exit(1);
    }
}

// Host SoA buffers for sld_atoms, gathered from and scattered to the AoS
// atom_struct array (struct.h:44-120).
struct Soa {
    double* mem;
    double *r, *p, *f, *s, *hh, *hl, *rho, *pe, *ke, *me, *me0, *stress, *vir, *sjs, *lv;
    explicit Soa(int n) {
        const size_t m = (size_t)n;
        mem = new double[m * (6 * 3 + 6 + 8)]();
        double* q = mem;
        double** vec3[] = {&r, &p, &f, &s, &hh, &hl};
        for (double** x : vec3) { *x = q; q += 3 * m; }
        stress = q; q += 6 * m;
        double** sc[] = {&rho, &pe, &ke, &me, &me0, &vir, &sjs, &lv};
        for (double** x : sc) { *x = q; q += m; }
    }
    ~Soa() { delete[] mem; }
    Soa(const Soa&) = delete;
    Soa& operator=(const Soa&) = delete;
    sld_atoms view() const {
        sld_atoms v = {r, p, f, s, hh, hl, rho, pe, ke, me, me0, stress, vir, sjs, lv};
        return v;
    }
};

void put3(double* d, int i, const ::vector& x) { d[3 * i] = x.x; d[3 * i + 1] = x.y; d[3 * i + 2] = x.z; }
::vector get3(const double* d, int i) { return vec_init(d[3 * i], d[3 * i + 1], d[3 * i + 2]); }

void gather(Soa& a) {
    for (int i = 0; i < natom; ++i) {
        const atom_struct* t = first_atom_ptr + i;
        put3(a.r, i, t->r); put3(a.p, i, t->p); put3(a.f, i, t->f);
#ifdef magmom
        put3(a.s, i, t->m);  // with magmom the library's spin field carries the moments
#else
        put3(a.s, i, t->s);
#endif
        put3(a.hh, i, t->Heff_H); put3(a.hl, i, t->Heff_L);
        a.rho[i] = t->rho; a.pe[i] = t->pe; a.ke[i] = t->ke; a.me[i] = t->me; a.me0[i] = t->me0;
        a.vir[i] = t->vir; a.sjs[i] = t->sum_Jij_sj; a.lv[i] = t->local_volume;
        const double st[6] = {t->stress11, t->stress22, t->stress33, t->stress12, t->stress23, t->stress31};
        for (int k = 0; k < 6; ++k) a.stress[6 * i + k] = st[k];
    }
}

void scatter(const Soa& a) {
    for (int i = 0; i < natom; ++i) {
        atom_struct* t = first_atom_ptr + i;
        t->r = get3(a.r, i); t->p = get3(a.p, i); t->f = get3(a.f, i);
        t->Heff_H = get3(a.hh, i); t->Heff_L = get3(a.hl, i);
        t->rho = a.rho[i]; t->pe = a.pe[i]; t->ke = a.ke[i]; t->me = a.me[i]; t->me0 = a.me0[i];
        t->vir = a.vir[i]; t->sum_Jij_sj = a.sjs[i];
#ifdef magmom
        // core_ds_CPU.cpp:37-45,73-81: the moments are the state, the spins follow
        t->m = get3(a.s, i);
        t->m0 = vec_length(t->m);
        t->s = vec_divide(t->m, -el_g);
        t->s0 = vec_length(t->s);
#else
        t->s = get3(a.s, i);
        t->s0 = vec_length(t->s);
        t->m = vec_times(-el_g, t->s);
        t->m0 = el_g * t->s0;
#endif
        t->stress11 = a.stress[6 * i]; t->stress22 = a.stress[6 * i + 1]; t->stress33 = a.stress[6 * i + 2];
        t->stress12 = a.stress[6 * i + 3]; t->stress23 = a.stress[6 * i + 4]; t->stress31 = a.stress[6 * i + 5];
        // the per-atom volume of the last force pass (calculate_force_energy_CPU.cpp:266-283)
        t->local_volume = a.lv[i];
    }
}

// the host box globals after a barostat moved the device box
// (scale_stress_CPU.cpp:74-84, scale_pressure_CPU.cpp:50-59)
void box_from_device() {
    double b[6];
    ok(sld_get_box(g_ctx, b), "sld_get_box");
    d.xx = b[0]; d.yx = b[1]; d.yy = b[2]; d.zx = b[3]; d.zy = b[4]; d.zz = b[5];
    Inv_d = inverse_box_vector(d);
    box_length.x = d.xx;
    box_length.y = sqrt(d.yx * d.yx + d.yy * d.yy);
    box_length.z = sqrt(d.zx * d.zx + d.zy * d.zy + d.zz * d.zz);
    box_length_half = vec_times(0.5e0, box_length);
    box_volume = d.xx * d.yy * d.zz;
    density = double(natom) / box_volume;
}

// one sld_check per printed step: the energy check (with the
// renormalizeEnergy feedback) and the temperature, stress, pressure and spin
// averages of that step, shared by the check_* functions that write them
const sld_obs& checked(int current_step) {
    if (g_obs_step != current_step) {
        ok(sld_set_total_time(g_ctx, total_time), "sld_set_total_time");
        ok(sld_check(g_ctx, current_step, &g_obs), "sld_check");
        g_obs_step = current_step;
    }
    return g_obs;
}

FILE* append(const char* front) {
    char name[256];
    std::snprintf(name, sizeof(name), "%s%s.dat", front, out_body);
    FILE* f = std::fopen(name, "a");
    if (!f) {
        std::fprintf(stderr, "libspilady_hip: cannot append to %s\n", name);
        std::exit(1);
    }
    return f;
}

}  // namespace

// ---- set-up and tear-down (initial_GPU.cu:36-102, copy_CPU_btw_GPU.cu:35-159,
// potential_GPU.cu:43-105)

void initial_GPU() {
    sld_params prm = {};
    prm.natom = natom;
    prm.interval_of_print_out = interval_of_print_out;
    const double box[6] = {d.xx, d.yx, d.yy, d.zx, d.zy, d.zz};
    for (int k = 0; k < 6; ++k) prm.box[k] = box[k];
    prm.step = step;
    prm.temperature = temperature;
    prm.atmass = atmass;
    prm.gamma_L_over_mass = gamma_L_over_mass;
    prm.gamma_S_HL = gamma_S_HL;
    prm.kappa_e = kappa_e;
    prm.rcut_pot = rcut_pot;
    prm.rcut_mag = rcut_mag;
    prm.rcut_max = rcut_max;
    prm.min_length_link_cell = min_length_link_cell;
    prm.start_time = start_time;
    prm.random_seed = (uint64_t)random_seed;
    prm.replica = 0;
    prm.displace_limit = displace_limit;
    prm.phi_limit = phi_limit;
    int on[kNSwitch] = {};
#ifdef spinlang
    on[kSpinlang] = 1;
#endif
#ifdef lattlang
    on[kLattlang] = 1;
#endif
#ifdef eltemp
    on[kEltemp] = 1;
#endif
#ifdef localcolmot
    on[kLocalcolmot] = 1;
#endif
#ifdef renormalizeEnergy
    on[kRenormalize] = 1;
#endif
#ifdef changestep
    on[kChangestep] = 1;
#endif
#ifdef extfield
    on[kExtfield] = 1;
#endif
#ifdef extforce
    on[kExtforce] = 1;
#endif
#ifdef quantumnoise
    on[kQuantumnoise] = 1;
    prm.msteps_quantum = Msteps_quantum;
    prm.nfrequency_quantum = Nfrequency_quantum;
#endif
#ifdef localvol
    on[kLocalvol] = 1;
    prm.rcut_vol = rcut_vol;
#endif
#ifdef magmom
    on[kMagmom] = 1;
#endif
    apply_switches(prm, on);
    Soa a(natom);
    gather(a);
    const sld_atoms v = a.view();
    // f and rho come from the input here; initialize() then calls
    // calculate_rho and calculate_force_energy unless readconf is on
    // (initial.cpp:78-83)
    ok(sld_create(&prm, &v, 0, &g_ctx), "sld_create");
#ifdef eltemp
    // map_cells set the cells' Te, uniform or from the readTe file
    // (map_cells.cpp:128-159)
    {
        double* te = new double[ncells];
        for (int i = 0; i < ncells; ++i) te[i] = (first_cell_ptr + i)->Te;
        ok(sld_set_cells_te(g_ctx, te), "sld_set_cells_te");
        delete[] te;
    }
#endif
#if defined extfield || defined extforce
    // external_field(-1) / external_force(-1) filled the host atoms
    {
        double* h = new double[3 * (size_t)natom]();
        double* f = new double[3 * (size_t)natom]();
        for (int i = 0; i < natom; ++i) {
            const atom_struct* t = first_atom_ptr + i;
            (void)t;
#ifdef extfield
            put3(h, i, t->Hext);
#endif
#ifdef extforce
            put3(f, i, t->fext);
#endif
        }
        const double* hp = nullptr;
        const double* fp = nullptr;
#ifdef extfield
        hp = h;
#endif
#ifdef extforce
        fp = f;
#endif
        ok(sld_set_external(g_ctx, hp, fp), "sld_set_external");
        delete[] h;
        delete[] f;
    }
#endif
#if defined STRESS || defined PRESSURE
    {
        sld_barostat b = {};
#ifdef STRESS
        b.stress_on = 1;
        b.interval_of_scale_stress = interval_of_scale_stress;
        b.stress[0] = stress_xx; b.stress[1] = stress_yx; b.stress[2] = stress_yy;
        b.stress[3] = stress_zx; b.stress[4] = stress_zy; b.stress[5] = stress_zz;
#endif
#ifdef PRESSURE
        b.pressure_on = 1;
        b.interval_of_scale_pressure = interval_of_scale_pressure;
        b.pressure = pressure;
#endif
        b.baro_damping_time = baro_damping_time;
        ok(sld_set_barostat(g_ctx, &b), "sld_set_barostat");
    }
#endif
}

// the tables are built and uploaded inside sld_create
void potential_table_GPU() {}
void free_potential_ptr_GPU() {}
void copy_CPU_to_GPU() {}

void copy_atoms_from_GPU_to_CPU() {
    Soa a(natom);
    sld_atoms v = a.view();
    ok(sld_download(g_ctx, &v), "sld_download");
    scatter(a);
}

void copy_cells_from_GPU_to_CPU() {
    double* m = new double[11 * (size_t)ncells]();
    sld_cells c = {m, m + ncells, m + 2 * ncells, m + 3 * ncells, m + 4 * ncells, m + 5 * ncells,
                   m + 6 * ncells, m + 7 * ncells, m + 8 * ncells, m + 9 * ncells, m + 10 * ncells};
    ok(sld_get_cells(g_ctx, &c), "sld_get_cells");
#ifdef eltemp
    for (int i = 0; i < ncells; ++i) {
        cell_struct* q = first_cell_ptr + i;
        q->Te = c.te[i]; q->Ee = c.ee[i]; q->Tl = c.tl[i]; q->Ts_R = c.ts_r[i]; q->Ts_L = c.ts_l[i];
        q->Gel = c.gel[i]; q->Ges = c.ges[i];
        q->sum_R_up = c.sum_r_up[i]; q->sum_R_dn = c.sum_r_dn[i];
        q->sum_L_up = c.sum_l_up[i]; q->sum_L_dn = c.sum_l_dn[i];
    }
#endif
    delete[] m;
}

void free_copy_CPU_to_GPU() {
    if (g_ctx) sld_destroy(g_ctx);
    g_ctx = nullptr;
}

// done inside sld_create (links_CPU.cpp:33-83, allocate_cells_CPU.cpp:33-125)
// and released by sld_destroy (allocate_cells_CPU.cpp:127-140)
void initial_links() {}
void allocate_cells() {}
void free_allocate_memory() {}
// the Langevin noise is a counter-based stream keyed by (random_seed, atom,
// pass): there is no generator state to create or free
// (random_number_CPU.cpp:38-113)
void initialize_random_number() {}
void free_random_number() {}
#ifdef quantumnoise
// the filter and the noise histories are built by sld_create and freed by
// sld_destroy (quantum_noise_CPU.cpp)
void initial_quantum_noise() {}
void free_quantum_noise() {}
#endif

// ---- the step (core.cpp:31-76)

void links() { ok(sld_links(g_ctx), "sld_links"); }
void calculate_rho() { ok(sld_calculate_rho(g_ctx), "sld_calculate_rho"); }
void calculate_force_energy() { ok(sld_calculate_force_energy(g_ctx), "sld_calculate_force_energy"); }
void core_ds(double dt) { ok(sld_core_ds(g_ctx, dt), "sld_core_ds"); }
void core_dp(double dt) { ok(sld_core_dp(g_ctx, dt), "sld_core_dp"); }
void core_dr(double dt) { ok(sld_core_dr(g_ctx, dt), "sld_core_dr"); }
#ifdef eltemp
void core_dTe(double dt) { ok(sld_core_dTe(g_ctx, dt), "sld_core_dTe"); }
void calculate_temperature() { ok(sld_calculate_temperature(g_ctx), "sld_calculate_temperature"); }
#endif

#ifdef extfield
// external_field_CPU.cpp:33-53: at step -1 the per-atom field of
// extfield.in, in tesla, converted to eV; it then stays as read (without the
// file the field stays zero: the reference's built-in time-dependent example
// is a template for users to edit)
void external_field(int current_step) {
    if (current_step != -1) return;
    ifstream infile("extfield.in");
    if (!infile) return;
    cout << "Reading external field file!!!" << '\n';
    int n = 0, temp = 0;
    infile >> n;
    for (int i = 0; i < n && i < natom; ++i) {
        atom_struct* t = first_atom_ptr + i;
        infile >> temp >> t->Hext.x >> t->Hext.y >> t->Hext.z;
        t->Hext = vec_times(-el_g * muB, t->Hext);
    }
}
#endif
#ifdef extforce
// external_force_CPU.cpp:33-50: the per-atom force of extforce.in (eV/A)
void external_force(int current_step) {
    if (current_step != -1) return;
    ifstream infile("extforce.in");
    if (!infile) return;
    cout << "Reading external forces file!!!" << '\n';
    int n = 0, temp = 0;
    infile >> n;
    for (int i = 0; i < n && i < natom; ++i) {
        atom_struct* t = first_atom_ptr + i;
        infile >> temp >> t->fext.x >> t->fext.y >> t->fext.z;
    }
}
#endif

// ---- observables, written as check_*_CPU.cpp write them (scientific, 15
// digits: the characters of %.15e)

// check_energy_CPU.cpp:33-140
void check_energy(int current_step) {
    const sld_obs& o = checked(current_step);
    FILE* f = append("enr-");
    std::fprintf(f, "%d %.15e %.15e %.15e %.15e %.15e", current_step, total_time, o.ave_pe, o.ave_ke, o.ave_me,
                 o.ave_me0);
#ifdef eltemp
    std::fprintf(f, " %.15e", o.ave_ee);
#endif
    std::fprintf(f, "\n");
    std::fclose(f);
}

// check_temperature_CPU.cpp:33-164
void check_temperature(int current_step) {
    const sld_obs& o = checked(current_step);
    FILE* f = append("tmp-");
    std::fprintf(f, "%d %.15e %.15e %.15e %.15e", current_step, total_time, o.tl, o.ts_r, o.ts_l);
#ifdef eltemp
    std::fprintf(f, " %.15e", o.te);
#endif
    std::fprintf(f, "\n");
    std::fclose(f);
}

// check_stress_CPU.cpp:33-83 (GPa)
void check_stress(int current_step) {
    const sld_obs& o = checked(current_step);
    const double* s = o.ave_stress;
    FILE* f = append("str-");
    std::fprintf(f, "%d %.15e %.15e %.15e %.15e %.15e %.15e %.15e %.15e %.15e %.15e %.15e %.15e %.15e %.15e %.15e\n",
                 current_step, total_time, d.xx, d.yx, d.yy, d.zx, d.zy, d.zz, density, (s[0] + s[1] + s[2]) / 3e0,
                 s[0], s[1], s[2], s[3], s[4], s[5]);
    std::fclose(f);
}

// check_pressure_CPU.cpp:33-60 (GPa)
void check_pressure(int current_step) {
    const sld_obs& o = checked(current_step);
    FILE* f = append("prs-");
    std::fprintf(f, "%d %.15e %.15e %.15e %.15e %.15e %.15e %.15e %.15e %.15e\n", current_step, total_time, d.xx,
                 d.yx, d.yy, d.zx, d.zy, d.zz, density, o.pressure);
    std::fclose(f);
}

// check_spin_CPU.cpp:33-81
void check_spin(int current_step) {
    const sld_obs& o = checked(current_step);
    FILE* f = append("spn-");
    std::fprintf(f, "%d %.15e %.15e %.15e %.15e %.15e %.15e %.15e %.15e %.15e\n", current_step, total_time,
                 o.ave_s[0], o.ave_s[1], o.ave_s[2], o.ave_s0, o.ave_m[0], o.ave_m[1], o.ave_m[2], o.ave_m0);
    std::fclose(f);
}

// ---- scale() (scale.cpp:31-44)

#ifdef changestep
// scale_step_CPU.cpp:33-97: the device adapts its time step; the host's
// follows it, since core() passes the global step to every entry point
void scale_step() {
    ok(sld_scale_step(g_ctx), "sld_scale_step");
    step = sld_get_step(g_ctx);
}
#endif
#ifdef STRESS
void scale_stress() {
    ok(sld_set_total_time(g_ctx, total_time), "sld_set_total_time");
    ok(sld_scale_stress(g_ctx), "sld_scale_stress");
    box_from_device();
}
#endif
#ifdef PRESSURE
void scale_pressure() {
    ok(sld_set_total_time(g_ctx, total_time), "sld_set_total_time");
    ok(sld_scale_pressure(g_ctx), "sld_scale_pressure");
    box_from_device();
}
#endif

#endif  // HIPLIB
