// spilady_hip: the SPILADY driver (spilady.cpp:33-72) on libspilady_hip.so.
//
//   spilady_hip [variables.in]
//
// Reads the run parameters from variables.in (read_variables.cpp:40-254; the
// control.h switches spinlang, lattlang, eltemp, localcolmot and
// renormalizeEnergy are runtime keys here, 1 = on), the initial configuration
// from the con- file named by in_config (read_config.cpp:32-147; the readconf
// path: f and rho are taken from the file, initial.cpp:78-83), then runs
// no_of_production_steps steps of core() (the runstep loop) with check() on
// every printed step and write() on every configuration step, appending to
// enr-/tmp-/spn-<out_body>.dat and writing con-/cel-<out_body>_NNNN.dat in
// the reference's formats (check_energy_CPU.cpp:113-137,
// check_temperature_CPU.cpp:136-161, check_spin_CPU.cpp:58-79,
// write_config.cpp:31-206).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/spilady_hip.h"

namespace {

// spilady.h:53-62
constexpr double kBoltz = 8.617343e-5;
constexpr double kSpeedOfLight = 299792458e0;
constexpr double kUtoEv = 931.494028e6 / kSpeedOfLight / kSpeedOfLight;
constexpr double kLengthScale = 1e-10;
constexpr double kElG = 2.0023193043622e0;   // electron g factor (spilady.h:53-62)
constexpr double kMuB = 5.7883817555e-5;     // eV/T

[[noreturn]] void fail(const std::string& msg) {
    std::fprintf(stderr, "spilady_hip: %s\n", msg.c_str());
    std::exit(1);
}

// variables.in: whitespace-separated "key value" pairs, later keys win; the
// defaults are the reference's (default.h:44-226)
struct Vars {
    std::map<std::string, std::string> kv = {
        {"OMP_threads", "2"}, {"out_body", "spilady"}, {"no_of_production_steps", "1000"}, {"step", "1e-15"},
        {"interval_of_print_out", "1"}, {"interval_of_config_out", "1000"}, {"in_config", "con-input.dat"},
        {"kappa_e", "80"}, {"start_time", "0"}, {"atmass", "55.847"}, {"temperature", "300"},
        {"gamma_L_over_mass", "6e11"}, {"gamma_S_HL", "5.88e13"}, {"random_seed", "1234"},
        {"rcut_pot", "4.1"}, {"rcut_mag", "3.75"}, {"rcut_max", "4.1"}, {"min_length_link_cell", "4.11"},
        {"spinlang", "1"}, {"lattlang", "1"}, {"eltemp", "1"}, {"localcolmot", "1"},
        {"renormalizeEnergy", "1"}, {"runstep", "1"}, {"changestep", "0"},
        {"total_production_time", "1e-12"}, {"displace_limit", "0.01"}, {"phi_limit", "0.628319"},
        {"extfield", "0"}, {"extforce", "0"},
        // control.h localvol with its cut-off 1.2 a_lattice (default.h:96,208-211), and magmom
        {"localvol", "0"}, {"rcut_vol", "3.396"}, {"magmom", "0"},
        // Berendsen barostats: control.h STRESS / PRESSURE, default.h:177-193
        {"STRESS", "0"}, {"PRESSURE", "0"}, {"interval_of_scale_stress", "1"}, {"stress_xx", "0"},
        {"stress_yx", "0"}, {"stress_yy", "0"}, {"stress_zx", "0"}, {"stress_zy", "0"}, {"stress_zz", "0"},
        {"interval_of_scale_pressure", "1"}, {"pressure", "0"}, {"baro_damping_time", "1e-13"},
        // quantum thermostat: control.h quantumnoise, default.h:222-225
        {"quantumnoise", "0"}, {"Msteps_quantum", "50"}, {"Nfrequency_quantum", "150"},
        // control.h readTe: the cells' electron temperatures from a cel- file (default.h:83-85)
        {"readTe", "0"}, {"in_eltemp", "cel-input.dat"},
        // control.h writevsim: V_sim files (default.h:67-69)
        {"writevsim", "0"}, {"interval_of_vsim", "1000"}, {"vsim_prec", "4"}};
    void read(const char* path) {
        std::ifstream in(path);
        if (!in) fail(std::string("cannot open ") + path);
        std::string k, v;
        while (in >> k >> v)
            if (kv.count(k)) kv[k] = v;
    }
    double d(const char* k) const { return std::atof(kv.at(k).c_str()); }
    long long i(const char* k) const { return (long long)std::atof(kv.at(k).c_str()); }
    const std::string& s(const char* k) const { return kv.at(k); }
};

struct Atoms {
    int n = 0;
    double total_time = 0.0;
    double box[6] = {0, 0, 0, 0, 0, 0};
    std::vector<double> r, p, f, s, hh, hl, stress, rho, pe, ke, me, me0, vir, sjs, lv;
    std::vector<std::string> el;  // element column of the con- file
    void resize(int natom) {
        n = natom;
        el.assign(n, "Fe");
        for (auto* v : {&r, &p, &f, &s, &hh, &hl}) v->assign((size_t)3 * n, 0.0);
        for (auto* v : {&rho, &pe, &ke, &me, &me0, &vir, &sjs, &lv}) v->assign(n, 0.0);
        stress.assign((size_t)6 * n, 0.0);
    }
    sld_atoms view() {
        return {r.data(), p.data(), f.data(), s.data(), hh.data(), hl.data(), rho.data(), pe.data(),
                ke.data(), me.data(), me0.data(), stress.data(), vir.data(), sjs.data(), lv.data()};
    }
};

// con- file (read_config.cpp:39-108): "natom total_time TYPE", two box lines,
// then "i element r p f stress[6] rho ke pe s s0 Heff_H Heff_L me me0 local_volume"
Atoms read_con(const std::string& path) {
    std::ifstream in(path);
    if (!in) fail("cannot open the initial configuration " + path);
    Atoms a;
    int n = 0;
    std::string type;
    in >> n >> a.total_time >> type;
    if (type != "SLDHL") fail("the input configuration file is for " + type + ", not for SLDHL");
    for (double& b : a.box) in >> b;
    a.resize(n);
    for (int i = 0; i < n; ++i) {
        int idx;
        double s0, hx, hy, hz, lx, ly, lz;
        in >> idx >> a.el[i];
        for (int k = 0; k < 3; ++k) in >> a.r[3 * i + k];
        for (int k = 0; k < 3; ++k) in >> a.p[3 * i + k];
        for (int k = 0; k < 3; ++k) in >> a.f[3 * i + k];
        for (int k = 0; k < 6; ++k) in >> a.stress[6 * i + k];
        in >> a.rho[i] >> a.ke[i] >> a.pe[i];
        for (int k = 0; k < 3; ++k) in >> a.s[3 * i + k];
        in >> s0 >> hx >> hy >> hz >> lx >> ly >> lz >> a.me[i] >> a.me0[i] >> a.lv[i];
        // the y columns repeat x (write_config.cpp:117,120); kept as read
        a.hh[3 * i] = hx; a.hh[3 * i + 1] = hy; a.hh[3 * i + 2] = hz;
        a.hl[3 * i] = lx; a.hl[3 * i + 1] = ly; a.hl[3 * i + 2] = lz;
        if (!in) fail("con- file ends before atom " + std::to_string(i));
    }
    return a;
}

std::string numbered(const char* front, const std::string& body, int n_output) {
    char buf[16];
    std::snprintf(buf, sizeof(buf), "_%04d.dat", n_output);
    return front + body + buf;
}

// check_energy/temperature/stress/pressure/spin (check.cpp:31-47): one line
// per printed step in each of enr-, tmp-, str-, prs- and spn-
// box: the box matrix at the check (it changes under the barostats)
void append_logs(const std::string& body, const sld_obs& o, const sld_params& prm, const double* box) {
    const int eltemp = prm.eltemp;
    FILE* e = std::fopen(("enr-" + body + ".dat").c_str(), "a");
    FILE* t = std::fopen(("tmp-" + body + ".dat").c_str(), "a");
    FILE* s = std::fopen(("spn-" + body + ".dat").c_str(), "a");
    FILE* st = std::fopen(("str-" + body + ".dat").c_str(), "a");
    FILE* pr = std::fopen(("prs-" + body + ".dat").c_str(), "a");
    if (!e || !t || !s || !st || !pr) fail("cannot append to the enr-/tmp-/spn-/str-/prs- files");
    // check_stress_CPU.cpp:68-81 and check_pressure_CPU.cpp:49-55: step, time,
    // the box matrix, the density, then the stresses or the pressure
    const double* b = box;
    const double density = (double)prm.natom / (b[0] * b[2] * b[5]);
    const double* sv = o.ave_stress;
    std::fprintf(st, "%d %.15e %.15e %.15e %.15e %.15e %.15e %.15e %.15e %.15e %.15e %.15e %.15e %.15e %.15e %.15e\n",
                 o.step, o.total_time, b[0], b[1], b[2], b[3], b[4], b[5], density, (sv[0] + sv[1] + sv[2]) / 3e0,
                 sv[0], sv[1], sv[2], sv[3], sv[4], sv[5]);
    std::fprintf(pr, "%d %.15e %.15e %.15e %.15e %.15e %.15e %.15e %.15e %.15e\n", o.step, o.total_time, b[0], b[1],
                 b[2], b[3], b[4], b[5], density, o.pressure);
    std::fclose(st);
    std::fclose(pr);
    std::fprintf(e, "%d %.15e %.15e %.15e %.15e %.15e", o.step, o.total_time, o.ave_pe, o.ave_ke, o.ave_me,
                 o.ave_me0);
    if (eltemp) std::fprintf(e, " %.15e", o.ave_ee);
    std::fprintf(e, "\n");
    std::fprintf(t, "%d %.15e %.15e %.15e %.15e", o.step, o.total_time, o.tl, o.ts_r, o.ts_l);
    if (eltemp) std::fprintf(t, " %.15e", o.te);
    std::fprintf(t, "\n");
    std::fprintf(s, "%d %.15e %.15e %.15e %.15e %.15e %.15e %.15e %.15e %.15e\n", o.step, o.total_time,
                 o.ave_s[0], o.ave_s[1], o.ave_s[2], o.ave_s0, o.ave_m[0], o.ave_m[1], o.ave_m[2], o.ave_m0);
    std::fclose(e);
    std::fclose(t);
    std::fclose(s);
}

// write_config.cpp:31-206
void write_config(sld_ctx* ctx, const std::string& body, int n_output, Atoms& a, int eltemp) {
    sld_atoms v = a.view();
    if (sld_download(ctx, &v)) fail(sld_last_error(ctx));
    if (sld_get_box(ctx, a.box)) fail(sld_last_error(ctx));
    const double t = sld_total_time(ctx);
    FILE* f = std::fopen(numbered("con-", body, n_output).c_str(), "w");
    if (!f) fail("cannot write the con- file");
    std::fprintf(f, "%d %.15e SLDHL\n", a.n, t);
    std::fprintf(f, "%.15e %.15e %.15e\n%.15e %.15e %.15e\n", a.box[0], a.box[1], a.box[2], a.box[3], a.box[4],
                 a.box[5]);
    // local_volume: box_volume / natom (calculate_force_energy_CPU.cpp:276;
    // box_volume = d.xx d.yy d.zz, read_config.cpp:126), or with localvol each
    // atom's own (:266-283), as sld_download gives it
    for (int i = 0; i < a.n; ++i) {
        const double* r = &a.r[3 * i];
        const double* p = &a.p[3 * i];
        const double* fo = &a.f[3 * i];
        const double* st = &a.stress[6 * i];
        const double* s = &a.s[3 * i];
        const double* hh = &a.hh[3 * i];
        const double* hl = &a.hl[3 * i];
        const double s0 = std::sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
        std::fprintf(f,
                     "%d %s %.15e %.15e %.15e %.15e %.15e %.15e %.15e %.15e %.15e %.15e %.15e %.15e %.15e %.15e "
                     "%.15e %.15e %.15e %.15e %.15e %.15e %.15e %.15e %.15e %.15e %.15e %.15e %.15e %.15e %.15e "
                     "%.15e %.15e\n",
                     i, a.el[i].c_str(), r[0], r[1], r[2], p[0], p[1], p[2], fo[0], fo[1], fo[2], st[0], st[1], st[2],
                     st[3], st[4],
                     st[5], a.rho[i], a.ke[i], a.pe[i], s[0], s[1], s[2], s0, hh[0], hh[0], hh[2], hl[0], hl[0],
                     hl[2], a.me[i], a.me0[i], a.lv[i]);
    }
    std::fclose(f);
    if (!eltemp) return;
    int n3[3], ng = 0;
    if (sld_get_geometry(ctx, n3, &ng, nullptr)) fail(sld_last_error(ctx));
    const int C = n3[0] * n3[1] * n3[2];
    std::vector<double> te(C), tl(C), tsr(C), tsl(C);
    sld_cells cv = {};
    cv.te = te.data(); cv.tl = tl.data(); cv.ts_r = tsr.data(); cv.ts_l = tsl.data();
    if (sld_get_cells(ctx, &cv)) fail(sld_last_error(ctx));
    FILE* g = std::fopen(numbered("cel-", body, n_output).c_str(), "w");
    if (!g) fail("cannot write the cel- file");
    const double* d = a.box;
    std::fprintf(g, "%d %.15e\n%.15e %.15e %.15e\n%.15e %.15e %.15e\n", C, t, d[0], d[1], d[2], d[3], d[4], d[5]);
    for (int ix = 0; ix < n3[0]; ++ix)
        for (int iy = 0; iy < n3[1]; ++iy)
            for (int iz = 0; iz < n3[2]; ++iz) {
                const int c = ix + iy * n3[0] + iz * n3[0] * n3[1];
                const double qx = (ix + 0.5) / n3[0], qy = (iy + 0.5) / n3[1], qz = (iz + 0.5) / n3[2];
                std::fprintf(g, "%.15e %.15e %.15e %.15e %.15e %.15e %.15e \n", d[0] * qx + d[1] * qy + d[3] * qz,
                             d[2] * qy + d[4] * qz, d[5] * qz, te[c] / kBoltz, tl[c], tsr[c], tsl[c]);
            }
    std::fclose(g);
}

}  // namespace

// write_vsim (write_vsim.cpp:33-134): V_sim files vsm-<body>_NNNN.ascii (the
// box, then "x y z element px py pz"), .spin ("i |s| theta phi" in degrees)
// and .dat (pe), scientific with vsim_prec digits
void write_vsim(sld_ctx* ctx, const std::string& body, int n_output, Atoms& a, int prec) {
    sld_atoms v = a.view();
    if (sld_download(ctx, &v)) fail(sld_last_error(ctx));
    if (sld_get_box(ctx, a.box)) fail(sld_last_error(ctx));
    const double t = sld_total_time(ctx);
    const std::string stem = numbered("vsm-", body, n_output);
    const std::string base = stem.substr(0, stem.size() - 4);  // without ".dat"
    std::ofstream fa(base + ".ascii"), fs(base + ".spin"), fc(base + ".dat");
    if (!fa || !fs || !fc) fail("cannot write the vsm- files");
    for (std::ofstream* o : {&fa, &fs, &fc}) *o << std::scientific << std::setprecision(prec);
    const double* d = a.box;
    fa << a.n << " " << t << '\n' << d[0] << " " << d[1] << " " << d[2] << '\n' << d[3] << " " << d[4] << " " << d[5]
       << '\n';
    fs << a.n << '\n';
    const double pi = 3.141592653589793238462e0;
    for (int i = 0; i < a.n; ++i) {
        const double* r = &a.r[3 * i];
        const double* p = &a.p[3 * i];
        const double* s = &a.s[3 * i];
        fa << r[0] << " " << r[1] << " " << r[2] << " " << a.el[i] << " " << p[0] << " " << p[1] << " " << p[2] << '\n';
        const double ss = std::sqrt(s[0] * s[0] + s[1] * s[1] + s[2] * s[2]);
        fs << i + 1 << " " << ss << " " << std::acos(s[2] / ss) * 180e0 / pi << " " << std::atan2(s[1], s[0]) * 180e0 / pi
           << '\n';
        fc << a.pe[i] << '\n';
    }
}

// variables.all (read_variables.cpp:278-418): the variables in effect, in
// the reference's order and with the stream's default formatting, for the
// switches this driver runs. OMP_threads is echoed as the reference's OpenMP
// build does.
void write_variables_all(const Vars& v, const sld_params& prm, const sld_barostat& baro, bool runstep,
                         bool read_te, bool vsim) {
    std::ofstream out("variables.all");
    if (!out) fail("cannot write variables.all");
    out << "OMP_threads " << v.i("OMP_threads") << '\n';
    out << "out_body " << v.s("out_body") << '\n';
    if (runstep) out << "no_of_production_steps " << v.i("no_of_production_steps") << '\n';
    else out << "total_production_time " << v.d("total_production_time") << '\n';
    out << "step " << prm.step << '\n';
    out << "interval_of_print_out " << prm.interval_of_print_out << '\n';
    out << "interval_of_config_out " << v.i("interval_of_config_out") << '\n';
    if (vsim) {
        out << "interval_of_vsim " << v.i("interval_of_vsim") << '\n';
        out << "vsim_prec " << v.i("vsim_prec") << '\n';
    }
    out << "in_config " << v.s("in_config") << '\n';
    if (read_te) out << "in_eltemp " << v.s("in_eltemp") << '\n';
    if (prm.eltemp) out << "kappa_e " << prm.kappa_e << '\n';
    out << "atmass " << prm.atmass / (kUtoEv * std::pow(kLengthScale, 2)) << '\n';
    if (prm.lattlang || prm.spinlang) out << "temperature " << prm.temperature / kBoltz << '\n';
    if (prm.lattlang) out << "gamma_L_over_mass " << prm.gamma_L_over_mass << '\n';
    if (prm.spinlang) out << "gamma_S_HL " << prm.gamma_S_HL << '\n';
    if (baro.stress_on) {
        out << "interval_of_scale_stress " << baro.interval_of_scale_stress << '\n';
        const char* names[6] = {"stress_xx", "stress_yx", "stress_yy", "stress_zx", "stress_zy", "stress_zz"};
        for (int k = 0; k < 6; ++k) out << names[k] << " " << baro.stress[k] << '\n';
    }
    if (baro.pressure_on) {
        out << "interval_of_scale_pressure " << baro.interval_of_scale_pressure << '\n';
        out << "pressure " << baro.pressure << '\n';
    }
    if (baro.stress_on || baro.pressure_on) out << "baro_damping_time " << baro.baro_damping_time << '\n';
    out << "random_seed " << prm.random_seed << '\n';
    out << "rcut_pot " << prm.rcut_pot << '\n';
    out << "rcut_mag " << prm.rcut_mag << '\n';
    out << "rcut_max " << prm.rcut_max << '\n';
    out << "min_length_link_cell " << prm.min_length_link_cell << '\n';
    if (prm.localvol) out << "rcut_vol " << prm.rcut_vol << '\n';  // read_variables.cpp:398-401
    if (prm.changestep) {
        out << "displace_limit " << prm.displace_limit << '\n';
        out << "phi_limit " << prm.phi_limit << '\n';
    }
    if (prm.quantumnoise) {
        out << "Msteps_quantum " << prm.msteps_quantum << '\n';
        out << "Nfrequency_quantum " << prm.nfrequency_quantum << '\n';
    }
}

int main(int argc, char** argv) {
    Vars v;
    v.read(argc > 1 ? argv[1] : "variables.in");
    Atoms a = read_con(v.s("in_config"));

    // read_variables.cpp:155-159,258-276: input units to internal units
    sld_params prm = {};
    prm.natom = a.n;
    prm.interval_of_print_out = (int)v.i("interval_of_print_out");
    for (int k = 0; k < 6; ++k) prm.box[k] = a.box[k];
    prm.step = v.d("step");
    prm.temperature = v.d("temperature") * kBoltz;
    prm.atmass = v.d("atmass") * kUtoEv * kLengthScale * kLengthScale;
    prm.gamma_L_over_mass = v.d("gamma_L_over_mass");
    prm.gamma_S_HL = v.d("gamma_S_HL");
    prm.kappa_e = v.d("kappa_e");
    prm.rcut_pot = v.d("rcut_pot");
    prm.rcut_mag = v.d("rcut_mag");
    prm.rcut_max = v.d("rcut_max");
    prm.min_length_link_cell = v.d("min_length_link_cell");
    if (prm.rcut_max > prm.min_length_link_cell) prm.min_length_link_cell = prm.rcut_max + 0.001;
    prm.start_time = v.d("start_time");
    prm.random_seed = (uint64_t)std::strtoull(v.s("random_seed").c_str(), nullptr, 10);
    prm.spinlang = (int)v.i("spinlang");
    prm.lattlang = (int)v.i("lattlang");
    prm.eltemp = (int)v.i("eltemp");
    prm.localcolmot = prm.eltemp ? (int)v.i("localcolmot") : 0;
    prm.renormalize_energy = prm.eltemp ? (int)v.i("renormalizeEnergy") : 0;
    prm.changestep = (int)v.i("changestep");
    prm.extfield = (int)v.i("extfield");
    prm.extforce = (int)v.i("extforce");
    prm.displace_limit = v.d("displace_limit");
    prm.phi_limit = v.d("phi_limit");
    prm.quantumnoise = (int)v.i("quantumnoise");
    prm.msteps_quantum = (int)v.i("Msteps_quantum");
    prm.nfrequency_quantum = (int)v.i("Nfrequency_quantum");
    prm.localvol = (int)v.i("localvol");
    prm.rcut_vol = v.d("rcut_vol");
    prm.magmom = (int)v.i("magmom");
    sld_barostat baro = {};
    baro.stress_on = (int)v.i("STRESS");
    baro.pressure_on = (int)v.i("PRESSURE");
    baro.interval_of_scale_stress = (int)v.i("interval_of_scale_stress");
    baro.interval_of_scale_pressure = (int)v.i("interval_of_scale_pressure");
    const char* stress_keys[6] = {"stress_xx", "stress_yx", "stress_yy", "stress_zx", "stress_zy", "stress_zz"};
    for (int k = 0; k < 6; ++k) baro.stress[k] = v.d(stress_keys[k]);
    baro.pressure = v.d("pressure");
    baro.baro_damping_time = v.d("baro_damping_time");
    const bool barostats = baro.stress_on || baro.pressure_on;
    const bool runstep = v.i("runstep") != 0;
    const bool vsim = v.i("writevsim") != 0;
    const int every_vsim = (int)std::max(1LL, v.i("interval_of_vsim"));
    const int vsim_prec = (int)v.i("vsim_prec");
    write_variables_all(v, prm, baro, runstep, prm.eltemp && v.i("readTe"), vsim);
    const double total_production_time = v.d("total_production_time");
    const std::string body = v.s("out_body");
    const long long nsteps = v.i("no_of_production_steps");
    const int every_config = (int)std::max(1LL, v.i("interval_of_config_out"));
    const int every_print = prm.interval_of_print_out > 0 ? prm.interval_of_print_out : 1;

    std::printf("SPILADY (libspilady_hip) starts: %d atoms\n", a.n);
    sld_ctx* ctx = nullptr;
    sld_atoms init = a.view();
    if (sld_create(&prm, &init, 0, &ctx)) fail(ctx ? sld_last_error(ctx) : "sld_create failed");

    // external_field(-1) / external_force(-1) (initial.cpp:58-63): per-atom
    // values from extfield.in (tesla) and extforce.in (eV/A) when present;
    // without a file the field or force stays zero
    if (prm.extfield || prm.extforce) {
        std::vector<double> hext, fext;
        auto read_ext = [&](const char* name, double scale, std::vector<double>& out) {
            std::ifstream in(name);
            if (!in) return false;
            int n = 0;
            in >> n;
            if (n != a.n) fail(std::string(name) + ": atom count differs from the configuration");
            out.assign((size_t)3 * n, 0.0);
            for (int i = 0; i < n; ++i) {
                int idx;
                double x, y, z;
                if (!(in >> idx >> x >> y >> z)) fail(std::string("cannot read ") + name);
                out[3 * i] = scale * x; out[3 * i + 1] = scale * y; out[3 * i + 2] = scale * z;
            }
            return true;
        };
        const double tesla_to_ev = -kElG * kMuB;  // external_field_CPU.cpp:50
        const bool hf = prm.extfield && read_ext("extfield.in", tesla_to_ev, hext);
        const bool ff = prm.extforce && read_ext("extforce.in", 1.0, fext);
        if (sld_set_external(ctx, hf ? hext.data() : nullptr, ff ? fext.data() : nullptr)) fail(sld_last_error(ctx));
        if (prm.eltemp && sld_calculate_temperature(ctx)) fail(sld_last_error(ctx));
    }

    if (barostats && sld_set_barostat(ctx, &baro)) fail(sld_last_error(ctx));

    // readTe (map_cells.cpp:33-60,125-160): "ncells total_time", two box lines,
    // then per cell, x outermost and z innermost, "x y z Te(K) Tl Ts_R Ts_L";
    // the cells take Te and Ee = E(Te). The file must describe this box.
    if (prm.eltemp && v.i("readTe")) {
        const std::string name = v.s("in_eltemp");
        std::ifstream in(name);
        if (!in) fail("cannot open " + name);
        int n3[3];
        if (sld_get_geometry(ctx, n3, nullptr, nullptr)) fail(sld_last_error(ctx));
        const int C = n3[0] * n3[1] * n3[2];
        int nc = 0;
        double t = 0.0, bx[6];
        if (!(in >> nc >> t)) fail("cannot read " + name);
        for (double& x : bx)
            if (!(in >> x)) fail("cannot read the box of " + name);
        if (nc != C) fail(name + ": the number of link cells differs from the configuration's");
        for (int k = 0; k < 6; ++k)
            if (std::fabs(bx[k] - prm.box[k]) > 1e-9 * (1.0 + std::fabs(prm.box[k])))
                fail(name + ": the box differs from the configuration's");
        std::vector<double> te(C);
        for (int ix = 0; ix < n3[0]; ++ix)
            for (int iy = 0; iy < n3[1]; ++iy)
                for (int iz = 0; iz < n3[2]; ++iz) {
                    double x, y, z, T, tl, tr, ts;
                    if (!(in >> x >> y >> z >> T >> tl >> tr >> ts)) fail("cannot read the cells of " + name);
                    te[ix + iy * n3[0] + iz * n3[0] * n3[1]] = T * kBoltz;
                }
        if (sld_set_cells_te(ctx, te.data())) fail(sld_last_error(ctx));
    }

    // check(-1) and write(-1) (spilady.cpp:38-40)
    sld_obs o;
    double box[6];
    if (sld_check(ctx, -1, &o)) fail(sld_last_error(ctx));
    if (sld_get_box(ctx, box)) fail(sld_last_error(ctx));
    append_logs(body, o, prm, box);
    write_config(ctx, body, 0, a, prm.eltemp);
    if (vsim) write_vsim(ctx, body, 0, a, vsim_prec);

    std::vector<sld_obs> obs;
    if (!runstep) {
        // the time loop (spilady.cpp:46-57): one step per call while
        // total_time <= total_production_time; dt may change every step
        // (changestep)
        obs.resize(1);
        long long s = 0;
        while (sld_total_time(ctx) <= total_production_time + 1e-18) {
            int nobs = 0;
            if (sld_get_box(ctx, box)) fail(sld_last_error(ctx));  // check() precedes scale()
            if (sld_step(ctx, (int)s, 1, obs.data(), &nobs)) fail(sld_last_error(ctx));
            if (nobs > 0) append_logs(body, obs[0], prm, box);
            std::printf("%lld total time = %.6e dt = %.3e\n", s, sld_total_time(ctx), sld_get_step(ctx));
            ++s;
            if (s % every_config == 0) write_config(ctx, body, (int)((s - 1) / every_config + 1), a, prm.eltemp);
            if (vsim && s % every_vsim == 0) write_vsim(ctx, body, (int)((s - 1) / every_vsim + 1), a, vsim_prec);
        }
        write_config(ctx, body, 9999, a, prm.eltemp);  // write(-2)
        if (vsim) write_vsim(ctx, body, 9999, a, vsim_prec);
        sld_destroy(ctx);
        std::printf("SPILADY (libspilady_hip) ends\n");
        return 0;
    }
    // the runstep loop (spilady.cpp:43-59): advance to the next configuration
    // output in one call; every printed step comes back as one record. With
    // a barostat the box may change after any step, so one step per call
    // and its box is logged with its check
    for (long long s = 0; s < nsteps;) {
        long long next = (s / every_config + 1) * every_config;
        if (vsim) next = std::min(next, (s / every_vsim + 1) * every_vsim);
        if (next > nsteps) next = nsteps;
        if (barostats) next = s + 1;
        const int n = (int)(next - s);
        obs.resize((size_t)n / every_print + 2);
        int nobs = 0;
        if (sld_get_box(ctx, box)) fail(sld_last_error(ctx));
        if (sld_step(ctx, (int)s, n, obs.data(), &nobs)) fail(sld_last_error(ctx));
        for (int k = 0; k < nobs; ++k) {
            append_logs(body, obs[k], prm, box);
            std::printf("%d total time = %.6e dt = %.3e\n", obs[k].step, obs[k].total_time, sld_get_step(ctx));
        }
        s = next;
        if (s % every_config == 0) write_config(ctx, body, (int)((s - 1) / every_config + 1), a, prm.eltemp);
        if (vsim && s % every_vsim == 0) write_vsim(ctx, body, (int)((s - 1) / every_vsim + 1), a, vsim_prec);
    }
    write_config(ctx, body, 9999, a, prm.eltemp);  // write(-2)
    if (vsim) write_vsim(ctx, body, 9999, a, vsim_prec);
    sld_destroy(ctx);
    std::printf("SPILADY (libspilady_hip) ends\n");
    return 0;
}
