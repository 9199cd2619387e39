// Host-side parameter handling: JSON defaults and parsing
// (MpsInputOutput::readInputFile, src/MpsInputOutput.cpp:58-641) and the
// derived constants of MpsParticle::setParameters
// (src/MpsParticle.cpp:136-227) with the bucket grid of
// MpsBucket::allocateMemory / calcDomainLimits (src/MpsBucket.cpp:20-295).
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "mps_gpu.h"
#include "mps_internal.h"

namespace mpsgpu {

// ------------------------------------------------------------- tiny JSON
// Objects, arrays, strings, numbers, literals; // and /* */ comments are
// whitespace; a repeated key keeps its last value.
struct JVal {
    enum Kind { NUL, BOOL, NUM, STR, ARR, OBJ } kind = NUL;
    bool b = false;
    double num = 0.0;
    std::string str;
    std::vector<JVal> arr;
    std::map<std::string, JVal> obj;
    const JVal* get(const char* k) const {
        if (kind != OBJ) return nullptr;
        auto it = obj.find(k);
        return it == obj.end() ? nullptr : &it->second;
    }
};

class JParser {
   public:
    explicit JParser(const std::string& s) : s_(s) {}
    bool parse(JVal& out, std::string& err) {
        try {
            value(out);
            ws();
            if (i_ != s_.size()) throw std::string("trailing characters");
        } catch (const std::string& e) {
            char buf[64];
            snprintf(buf, sizeof(buf), " at byte %zu", i_);
            err = e + buf;
            return false;
        }
        return true;
    }

   private:
    const std::string& s_;
    size_t i_ = 0;
    void ws() {
        for (;;) {
            while (i_ < s_.size() && isspace((unsigned char)s_[i_])) i_++;
            if (i_ + 1 < s_.size() && s_[i_] == '/' && s_[i_ + 1] == '/') {
                while (i_ < s_.size() && s_[i_] != '\n') i_++;
            } else if (i_ + 1 < s_.size() && s_[i_] == '/' && s_[i_ + 1] == '*') {
                size_t e = s_.find("*/", i_ + 2);
                if (e == std::string::npos) throw std::string("unterminated comment");
                i_ = e + 2;
            } else {
                return;
            }
        }
    }
    char peek() {
        ws();
        if (i_ >= s_.size()) throw std::string("unexpected end");
        return s_[i_];
    }
    void expect(char c) {
        if (peek() != c) throw std::string("expected '") + c + "'";
        i_++;
    }
    void value(JVal& v) {
        char c = peek();
        if (c == '{') {
            v.kind = JVal::OBJ;
            i_++;
            if (peek() == '}') { i_++; return; }
            for (;;) {
                JVal k;
                if (peek() != '"') throw std::string("expected key");
                string(k.str);
                expect(':');
                JVal val;
                value(val);
                v.obj[k.str] = std::move(val);
                char d = peek();
                i_++;
                if (d == '}') return;
                if (d != ',') throw std::string("expected ',' or '}'");
            }
        } else if (c == '[') {
            v.kind = JVal::ARR;
            i_++;
            if (peek() == ']') { i_++; return; }
            for (;;) {
                JVal val;
                value(val);
                v.arr.push_back(std::move(val));
                char d = peek();
                i_++;
                if (d == ']') return;
                if (d != ',') throw std::string("expected ',' or ']'");
            }
        } else if (c == '"') {
            v.kind = JVal::STR;
            string(v.str);
        } else if (s_.compare(i_, 4, "true") == 0) {
            v.kind = JVal::BOOL; v.b = true; i_ += 4;
        } else if (s_.compare(i_, 5, "false") == 0) {
            v.kind = JVal::BOOL; v.b = false; i_ += 5;
        } else if (s_.compare(i_, 4, "null") == 0) {
            v.kind = JVal::NUL; i_ += 4;
        } else {
            const char* p = s_.c_str() + i_;
            char* e = nullptr;
            v.num = strtod(p, &e);
            if (e == p) throw std::string("bad value");
            v.kind = JVal::NUM;
            i_ += (size_t)(e - p);
        }
    }
    void string(std::string& out) {
        expect('"');
        out.clear();
        while (i_ < s_.size() && s_[i_] != '"') {
            if (s_[i_] == '\\' && i_ + 1 < s_.size()) {
                char e = s_[i_ + 1];
                out += (e == 'n') ? '\n' : (e == 't') ? '\t' : e;
                i_ += 2;
            } else {
                out += s_[i_++];
            }
        }
        if (i_ >= s_.size()) throw std::string("unterminated string");
        i_++;
    }
};

// json.value(key, default) semantics: a missing key (or a type that does not
// convert) keeps the default.
static double jnum(const JVal* o, const char* k, double def) {
    const JVal* v = o ? o->get(k) : nullptr;
    if (!v) return def;
    if (v->kind == JVal::NUM) return v->num;
    if (v->kind == JVal::BOOL) return v->b ? 1.0 : 0.0;
    return def;
}
static int jint(const JVal* o, const char* k, int def) {
    const JVal* v = o ? o->get(k) : nullptr;
    if (!v) return def;
    if (v->kind == JVal::NUM) return (int)v->num;
    if (v->kind == JVal::BOOL) return v->b ? 1 : 0;
    return def;
}
static std::string jstr(const JVal* o, const char* k, const char* def) {
    const JVal* v = o ? o->get(k) : nullptr;
    if (!v || v->kind != JVal::STR) return def;
    return v->str;
}

void paramsDefaults(mps_params* p) {
    memset(p, 0, sizeof(*p));
    // defaults as in src/MpsInputOutput.cpp:136-500
    p->wall_type = 1; p->fem_on = 0; p->forced_on = 0;
    p->domain_type_bc = 0; p->limit_type_bc = 0; p->num_bc = 1;
    p->density_fluid = 1000.0; p->density_wall = 1000.0; p->knm_vs1 = 0.000001; p->fluid_type = 0;
    p->gravity[0] = 0.0; p->gravity[1] = 0.0; p->gravity[2] = -9.81;
    p->knm_vs2 = 0.000001; p->dns_fl1 = 1000.0; p->dns_fl2 = 1540.0;
    p->dns_sdt = 1540.0; p->power_n = 1.2; p->meu0 = 0.03;
    p->phi_1 = 0.541; p->phi_wal = 0.541; p->phi_bed = 0.541; p->phi_2 = 0.6;
    p->cohes = 0.0; p->fraction_method = 2; p->dg = 0.0035; p->i0 = 0.75; p->mm = 100.0; p->stress_calc_method = 1;
    p->cd = 0.47; p->vf_min = 0.25; p->vf_max = 0.65;
    p->dim = 3.0; p->part_dist = 0.01; p->time_step = 0.0005; p->time_simulation = 1.0;
    p->iter_output = 80; p->cfl_number = 0.2; p->weight_type = 0; p->slip_condition = 0;
    p->re_s = 2.1; p->re_l = 2.1;
    p->gradient_type = 0; p->gradient_correction = 0; p->relax_press = 1.0;
    p->mps_type = 1; p->sound_speed = 15.0; p->gamma = 7.0;
    p->solver_type = 0; p->alpha_compressibility = 0.000001; p->relax_pnd = 0.001;
    p->shifting_type = 2; p->dri = 0.01; p->coef_a = 2.0; p->mach_number = 0.1;
    p->pnd_type = 0; p->diffusive_coef = 0.35;
    p->repulsive_force_type = 2; p->re_repulsive_force = 0.5; p->expect_max_velocity = 6.0;
    p->rep_coef_mitsume = 40000000.0; p->rep_coef_lennard_jones = 2.0; p->rep_coef_monaghan_kajtar = 1.0;
    p->eps_re = 0.01;
    p->free_surf_type = 0; p->pnd_threshold = 0.98; p->neigh_threshold = 0.85; p->npcd_threshold = 0.20;
    p->theta_threshold = 45.0; p->norm_threshold = 0.1;
    p->collision_type = 0; p->collision_ratio = 0.20; p->dist_limit_ratio = 0.85; p->lambda_collision = 0.20;
    p->ghost = -1; p->fluid = 0; p->wall = 2; p->dummy_wall = 3;
    p->surface = 1; p->inner = 0; p->other = -1;
    // extensions
    p->solver_tol = 0.0;       // Eigen default (DBL_EPSILON)
    p->solver_max_iter = 0;    // Eigen default (2 N)
    p->precision_mode = MPS_PREC_FP64;
    p->shift_semantics = MPS_SHIFT_JACOBI;
    p->ppe_transpose = 1;
    p->mixed_inner_tol = 1e-4;
    p->learned_mode = 0;
    p->ppe_warm_start = 0;
    p->forced_motion = 1;
    p->forced_center[0] = 0.45;
    p->forced_center[1] = 0.0;
    p->forced_theta_max = 4.0;
    p->forced_period = 1.9171;
    p->forced_velocity = 0.0;
    p->uniform_vel_wall[0] = p->uniform_vel_wall[1] = p->uniform_vel_wall[2] = 0.0;
}

int paramsFromJson(const char* path, mps_params* p, std::string& grid, std::string& mesh, std::string& err) {
    std::ifstream f(path);
    if (!f) { err = std::string("cannot open ") + path; return MPS_E_IO; }
    std::stringstream ss;
    ss << f.rdbuf();
    std::string text = ss.str();
    JVal root;
    if (!JParser(text).parse(root, err)) { err = "JSON parse error: " + err; return MPS_E_IO; }
    paramsDefaults(p);
    const JVal* fl = root.get("flags");
    p->wall_type = jint(fl, "wall_type", 1);
    p->fem_on = jint(fl, "fem_MESH", 0);
    p->forced_on = jint(fl, "forced_MESH", 0);
    const JVal* pn = root.get("pathNames");
    grid = jstr(pn, "particle_grid_file", "oops");
    mesh = jstr(pn, "mesh_rigid_file", "oops");
    // forced mesh motion: no reference key (the reference hard-codes the law)
    if (const JVal* fm = root.get("forced_motion")) {
        p->forced_motion = jint(fm, "rotation", p->forced_motion);
        p->forced_center[0] = jnum(fm, "center_x", p->forced_center[0]);
        p->forced_center[1] = jnum(fm, "center_z", p->forced_center[1]);
        p->forced_theta_max = jnum(fm, "theta_max_deg", p->forced_theta_max);
        p->forced_period = jnum(fm, "period", p->forced_period);
        p->forced_velocity = jnum(fm, "velocity", p->forced_velocity);
    }
    const JVal* dom = root.get("domain");
    const JVal* dmin = dom ? dom->get("min") : nullptr;
    const JVal* dmax = dom ? dom->get("max") : nullptr;
    const char* ax[3] = {"x", "y", "z"};
    for (int k = 0; k < 3; k++) {
        p->domain_min[k] = jnum(dmin, ax[k], 0.0);
        p->domain_max[k] = jnum(dmax, ax[k], 0.0);
    }
    const JVal* bnd = dom ? dom->get("boundary") : nullptr;
    if (bnd) {
        p->domain_type_bc = jint(bnd, "type", 0);
        p->limit_type_bc = jint(bnd, "limit", 0);
        p->num_bc = 1;
        const JVal* dir = bnd->get("direction");
        for (int k = 0; k < 3; k++) p->periodic_dir[k] = jint(dir, ax[k], 0);
    }
    const JVal* ph = root.get("physical");
    p->density_fluid = jnum(ph, "fluid_density", 1000.0);
    p->density_wall = jnum(ph, "wall_density", 1000.0);
    p->knm_vs1 = jnum(ph, "kinematic_visc", 0.000001);
    p->fluid_type = jint(ph, "fluid_type", 0);
    const JVal* gr = ph ? ph->get("gravity") : nullptr;
    p->gravity[0] = jnum(gr, "x", 0.0);
    p->gravity[1] = jnum(gr, "y", 0.0);
    p->gravity[2] = jnum(gr, "z", -9.81);
    const JVal* rh = ph ? ph->get("rheological") : nullptr;
    p->knm_vs2 = jnum(rh, "kinematic_visc_phase_2", 0.000001);
    p->dns_fl1 = jnum(rh, "fluid_density_phase_1", 1000.0);
    p->dns_fl2 = jnum(rh, "fluid_density_phase_2", 1540.0);
    p->dns_sdt = jnum(rh, "sediment_density", 1540.0);
    p->power_n = jnum(rh, "power_law_index", 1.2);
    p->meu0 = jnum(rh, "consistency_index", 0.03);
    const JVal* phi = rh ? rh->get("phi") : nullptr;
    p->phi_1 = jnum(phi, "lower", 0.541);
    p->phi_wal = jnum(phi, "wall", 0.541);
    p->phi_bed = jnum(phi, "bed", 0.541);
    p->phi_2 = jnum(phi, "second", 0.6);
    p->cohes = jnum(rh, "cohes_coeff", 0.0);
    p->fraction_method = jint(rh, "fraction_method", 2);
    p->dg = jnum(rh, "grain_size", 0.0035);
    p->i0 = jnum(rh, "I0", 0.75);
    p->mm = jnum(rh, "mm", 100.0);
    p->stress_calc_method = jint(rh, "stress_calc_method", 1);
    p->cd = jnum(rh, "drag_coeff", 0.47);
    const JVal* vfr = rh ? rh->get("volume_fraction") : nullptr;
    p->vf_min = jnum(vfr, "min", 0.25);
    p->vf_max = jnum(vfr, "max", 0.65);
    const JVal* nu = root.get("numerical");
    p->dim = jnum(nu, "dimension", 3.0);
    p->part_dist = jnum(nu, "particle_dist", 0.01);
    p->time_step = jnum(nu, "time_step", 0.0005);
    p->time_simulation = jnum(nu, "final_time", 1.0);
    p->iter_output = jint(nu, "iter_output", 80);
    p->cfl_number = jnum(nu, "CFL_number", 0.2);
    p->weight_type = jint(nu, "weight_type", 0);
    p->slip_condition = jint(nu, "slip_condition", 0);
    const JVal* re = nu ? nu->get("effective_radius") : nullptr;
    p->re_s = jnum(re, "small", 2.1);
    p->re_l = jnum(re, "large", 2.1);
    const JVal* gd = nu ? nu->get("gradient") : nullptr;
    p->gradient_type = jint(gd, "type", 0);
    p->gradient_correction = jint(gd, "correction", 0);
    p->relax_press = jnum(gd, "relax_fact", 1.0);
    p->mps_type = jint(nu, "mps_type", 1);
    const JVal* ex = nu ? nu->get("explicit_mps") : nullptr;
    const JVal* eq = ex ? ex->get("equation_state") : nullptr;
    p->sound_speed = jnum(eq, "speed_sound", 15.0);
    p->gamma = jnum(eq, "gamma", 7.0);
    const JVal* si = nu ? nu->get("semi_implicit_mps") : nullptr;
    p->solver_type = jint(si, "solver_type", 0);
    const JVal* wc = si ? si->get("weak_compressibility") : nullptr;
    p->alpha_compressibility = jnum(wc, "alpha", 0.000001);
    const JVal* st = si ? si->get("source_term") : nullptr;
    p->relax_pnd = jnum(st, "relax_pnd", 0.001);
    // extension keys (no reference counterpart)
    p->solver_tol = jnum(si, "tolerance", 0.0);
    p->solver_max_iter = jint(si, "max_iterations", 0);
    p->precision_mode = jint(si, "precision_mode", MPS_PREC_FP64);
    p->ppe_transpose = jint(si, "eigen_transpose", 1);
    p->mixed_inner_tol = jnum(si, "mixed_inner_tolerance", 1e-4);
    const JVal* ln = si ? si->get("learned") : nullptr;
    p->learned_mode = jint(ln, "mode", 0);
    p->ppe_warm_start = jint(si, "warm_start", 0);
    const JVal* sh = nu ? nu->get("particle_shifting") : nullptr;
    p->shifting_type = jint(sh, "type", 2);
    p->dri = jnum(sh, "DRI", 0.01);
    p->coef_a = jnum(sh, "coef_A", 2.0);
    p->mach_number = jnum(sh, "mach_number", 0.1);
    p->shift_semantics = jint(sh, "semantics", MPS_SHIFT_JACOBI);
    const JVal* pd = nu ? nu->get("pnd") : nullptr;
    p->pnd_type = jint(pd, "type", 0);
    p->diffusive_coef = jnum(pd, "diffusive_coeff", 0.35);
    const JVal* wf = nu ? nu->get("wall_repulsive_force") : nullptr;
    p->repulsive_force_type = jint(wf, "type", 2);
    p->re_repulsive_force = jnum(wf, "re", 0.5);
    p->expect_max_velocity = jnum(wf, "maxVel", 6.0);
    const JVal* wfc = wf ? wf->get("coefficient") : nullptr;
    p->rep_coef_mitsume = jnum(wfc, "Mitsume", 40000000.0);
    p->rep_coef_lennard_jones = jnum(wfc, "Lennard-Jones", 2.0);
    p->rep_coef_monaghan_kajtar = jnum(wfc, "Monaghan-Kajtar", 1.0);
    p->eps_re = jnum(wf, "eps_re", 0.01);
    const JVal* fs = nu ? nu->get("free_surface_threshold") : nullptr;
    p->free_surf_type = jint(fs, "type", 0);
    p->pnd_threshold = jnum(fs, "pnd", 0.98);
    p->neigh_threshold = jnum(fs, "neigh", 0.85);
    p->npcd_threshold = jnum(fs, "NPCD", 0.20);
    p->theta_threshold = jnum(fs, "ARC", 45.0);
    p->norm_threshold = jnum(fs, "normal", 0.1);
    const JVal* co = nu ? nu->get("particle_collision") : nullptr;
    p->collision_type = jint(co, "type", 0);
    p->collision_ratio = jnum(co, "ratio", 0.20);
    p->dist_limit_ratio = jnum(co, "dist_limit_ratio", 0.85);
    p->lambda_collision = jnum(co, "lambda", 0.20);
    const JVal* pt = nu ? nu->get("particle_type") : nullptr;
    p->ghost = jint(pt, "ghost", -1);
    p->fluid = jint(pt, "fluid", 0);
    p->wall = jint(pt, "wall", 2);
    p->dummy_wall = jint(pt, "dummyWall", 3);
    const JVal* bt = nu ? nu->get("boundary_type") : nullptr;
    p->surface = jint(bt, "free_surface", 1);
    p->inner = jint(bt, "inner", 0);
    p->other = jint(bt, "other", -1);
    if (p->part_dist * p->part_dist <= 10.0 * DBL_EPSILON) {
        err = "particle distance too small";
        return MPS_E_ARG;
    }
    if ((p->mps_type == MPS_IMPLICIT_PND || p->mps_type == MPS_IMPLICIT_PND_DIVU) && p->pnd_type != MPS_PND_SUM_WIJ) {
        err = "set pnd.type to 0 for incompressible MPS";   // src/MpsInputOutput.cpp:539-544
        return MPS_E_ARG;
    }
    return MPS_OK;
}

static inline double wfun(double dst, double re, int t) {
    switch (t) {
        case 0: return re / dst - 1.0;
        case 1: return re / dst + dst / re - 2.0;
        case 2: return re / dst - dst / re;
        case 3: return (1.0 - dst / re) * (1.0 - dst / re) * (1.0 - dst / re);
        case 4: return (1.0 - dst / re) * (1.0 - dst / re);
        default: return re / dst - 1.0;
    }
}

int paramsDerive(mps_params* p, int n, const double* pos, std::string& err) {
    if (p->derived) return MPS_OK;
    const double eps0 = 10.0 * DBL_EPSILON;
    const double l0 = p->part_dist;
    p->eps_zero = eps0;
    const bool periodic = p->domain_type_bc == 1;
    int pdir = -1;
    if (p->domain_type_bc != 0 && !periodic) { err = "unknown domain boundary type"; return MPS_E_UNSUPPORTED; }
    if (periodic) {
        // MpsBucket::calcDomainLimits (src/MpsBucket.cpp:141-287) handles one
        // periodic direction, with particle or polygon walls
        const int nper = (p->periodic_dir[0] != 0) + (p->periodic_dir[1] != 0) + (p->periodic_dir[2] != 0);
        if (nper != 1) { err = "periodic domains need exactly one periodic direction"; return MPS_E_UNSUPPORTED; }
        for (int k = 0; k < 3; k++) if (p->periodic_dir[k]) pdir = k;
        if ((int)p->dim == 2 && pdir == 2) { err = "periodic z in a 2-D case"; return MPS_E_UNSUPPORTED; }
    }
    if (p->fluid_type != 0 && p->fluid_type != 1) { err = "unknown fluid type"; return MPS_E_UNSUPPORTED; }
    if (p->fluid_type == 1 && p->wall_type == MPS_WALL_POLYGON && (int)p->dim == 2) {
        // the polygon-wall rheology indexes its bed grid with a 3-D column
        // (src/MpsViscosity.cpp:766-783)
        err = "the granular phase with polygon walls needs a 3-D case";
        return MPS_E_UNSUPPORTED;
    }
    if ((p->fem_on || p->forced_on) && p->wall_type != MPS_WALL_POLYGON) {
        err = "deformable / forced meshes need polygon walls";
        return MPS_E_UNSUPPORTED;
    }
    if (p->forced_on && p->forced_motion == 1 && !(p->forced_period > 0.0)) {
        err = "forced rotation needs a positive period";
        return MPS_E_ARG;
    }
    // domain extension (setParameters :139-148)
    for (int k = 0; k < 3; k++) { p->domain_min[k] -= l0 * 3.0; p->domain_max[k] += l0 * 3.0; }
    if ((int)p->dim == 2) { p->domain_min[2] = 0.0; p->domain_max[2] = 0.0; }
    // MpsParticle::setParameters (src/MpsParticle.cpp:151-155)
    p->num_of_meshs = (p->wall_type == 1 ? 1 : 0) + (p->fem_on ? 1 : 0) + (p->forced_on ? 1 : 0);
    p->re_s = l0 * p->re_s;
    p->re_l = l0 * p->re_l;
    p->re_repulsive_force = l0 * p->re_repulsive_force;
    p->re_s2 = p->re_s * p->re_s;
    p->re_l2 = p->re_l * p->re_l;
    p->eps_re = p->eps_re * p->re_s2 / 4.0;
    // lattice sums over a full compact support (:166-195)
    double n0S = 0.0, n0L = 0.0, n0G = 0.0, lam = 0.0, cnt = 0.0;
    const int lo = (int)ceil(p->re_l / l0) + 1, hi = (int)ceil(p->re_l / l0) + 2;
    const bool two = (int)p->dim == 2;
    const int zlo = two ? 0 : -lo, zhi = two ? 1 : hi;
    for (int ix = -lo; ix < hi; ix++)
        for (int iy = -lo; iy < hi; iy++)
            for (int iz = zlo; iz < zhi; iz++) {
                const double x = l0 * (double)ix, y = l0 * (double)iy, z = l0 * (double)iz;
                const double d2 = x * x + y * y + z * z;
                if (d2 > p->re_l2 || d2 <= eps0) continue;
                const double d = sqrt(d2);
                n0L += wfun(d, p->re_l, p->weight_type);
                lam += d2 * wfun(d, p->re_l, p->weight_type);
                cnt += 1;
                if (d2 <= p->re_s2) {
                    n0S += wfun(d, p->re_s, p->weight_type);
                    n0G += wfun(d, p->re_s, p->weight_type);
                }
            }
    p->pnd_small_zero = n0S;
    p->pnd_large_zero = n0L;
    p->pnd_gradient_zero = n0G;
    p->num_neigh_zero = cnt;
    p->lambda_zero = lam / n0L;
    p->coeff_viscosity = 2.0 * p->knm_vs1 * p->dim / (n0L * p->lambda_zero);
    p->coeff_visc_multiphase = 2.0 * p->dim / (n0L * p->lambda_zero);
    p->coeff_press_emps = p->sound_speed * p->sound_speed / n0S;
    p->coeff_press_grad = -p->dim / n0G;
    p->coeff_press_wcmps = p->sound_speed * p->sound_speed;
    p->coeff_shifting1 = p->dri * l0 / n0S;
    p->coeff_shifting2 = p->coef_a * l0 * l0 * p->cfl_number * p->mach_number;
    p->coeff_ppe = 2.0 * p->dim / (n0L * p->lambda_zero);
    p->coeff_ppe_source = p->relax_pnd / (p->time_step * p->time_step * n0S);
    p->dns_fluid = p->density_fluid;
    p->dns_wall = p->density_wall;
    p->inv_dns_fluid = 1.0 / p->density_fluid;
    p->inv_dns_wall = 1.0 / p->density_wall;
    p->inv_part_dist = 1.0 / l0;
    p->dist_collision_limit = l0 * p->dist_limit_ratio;
    p->dist_collision_limit2 = p->dist_collision_limit * p->dist_collision_limit;
    p->restitution_collision = 1.0 + p->collision_ratio;
    p->beta_pnd = p->pnd_threshold * n0S;
    p->beta_neigh = p->neigh_threshold * cnt;
    p->delta2 = p->npcd_threshold * p->npcd_threshold * l0 * l0;
    p->theta_arc = p->theta_threshold / 180.0 * 3.14159265;
    p->h_threshold2 = 1.33 * 1.33 * l0 * l0;
    p->dst_threshold2 = 2.0 * p->h_threshold2;
    p->norm_threshold2 = p->norm_threshold * p->norm_threshold;
    // bucket grid (allocateMemory :20-67)
    p->bucket_side = p->re_l * (1.0 + p->cfl_number);
    p->inv_bucket_side = 1.0 / p->bucket_side;
    p->num_buckets_x = (int)((p->domain_max[0] - p->domain_min[0]) * p->inv_bucket_side) + 3;
    p->num_buckets_y = (int)((p->domain_max[1] - p->domain_min[1]) * p->inv_bucket_side) + 3;
    p->num_buckets_z = (int)((p->domain_max[2] - p->domain_min[2]) * p->inv_bucket_side) + 3;
    if (two) p->num_buckets_z = 1;
    for (int k = 0; k < 3; k++) p->periodic_length[k] = 0.0;
    if (p->wall_type == MPS_WALL_PARTICLE || periodic) {
        // calcDomainLimits (:71-295), called for particle walls or a periodic
        // domain (src/MpsBucket.cpp:45-48)
        double lim[3][2];
        if (p->limit_type_bc == 0) {
            if (n <= 0 || !pos) { err = "particle walls need particle positions to size the domain"; return MPS_E_ARG; }
            for (int k = 0; k < 3; k++) lim[k][0] = lim[k][1] = pos[k];
            for (int i = 0; i < n; i++)
                for (int k = 0; k < 3; k++) {
                    lim[k][0] = std::min(lim[k][0], pos[i * 3 + k]);
                    lim[k][1] = std::max(lim[k][1], pos[i * 3 + k]);
                }
            int td = 0;
            for (int k = 0; k < 3; k++) td += (lim[k][0] != lim[k][1]);
            if (td != (int)p->dim) { err = "dimensions in json and grid do not match"; return MPS_E_ARG; }
        } else {
            for (int k = 0; k < 3; k++) { lim[k][0] = p->domain_min[k]; lim[k][1] = p->domain_max[k]; }
        }
        int* nbk[3] = {&p->num_buckets_x, &p->num_buckets_y, &p->num_buckets_z};
        if (!periodic) {
            const int nd = two ? 2 : 3;
            for (int k = 0; k < nd; k++) { lim[k][0] -= 0.5 * l0; lim[k][1] += 0.5 * l0; }
            for (int k = 0; k < (int)p->dim; k++) {
                lim[k][0] -= p->bucket_side;
                lim[k][1] += p->bucket_side;
                long nb = (long)((lim[k][1] - lim[k][0]) / p->bucket_side + 1);
                *nbk[k] = (int)nb;
                lim[k][1] = lim[k][0] + nb * p->bucket_side;
            }
        } else {
            // periodic direction d (:148-160 for x, :194-206 for y, :240-252 for z)
            const int d = pdir;
            p->periodic_length[d] = lim[d][1] - lim[d][0] + l0;
            lim[d][0] -= l0 * 0.5;
            lim[d][1] += l0 * 0.5;
            p->phys_min[d] = lim[d][0];
            p->phys_max[d] = lim[d][1];
            long nbd = (long)((lim[d][1] - lim[d][0]) / p->bucket_side);
            // the bucket side divides the periodic length exactly
            p->bucket_side = p->periodic_length[d] / (long)(p->periodic_length[d] / p->bucket_side);
            lim[d][0] -= p->bucket_side;
            lim[d][1] += p->bucket_side;
            nbd += 2;
            *nbk[d] = (int)nbd;
            // the other axes: one bucket beyond the wall particles, or beyond
            // the domain of the JSON file with polygon walls
            for (int k = 0; k < 3; k++) {
                if (k == d || (k == 2 && two)) continue;
                if (p->wall_type == MPS_WALL_PARTICLE) {
                    lim[k][0] -= p->bucket_side;
                    lim[k][1] += p->bucket_side;
                } else {
                    lim[k][0] = p->domain_min[k] - p->bucket_side;
                    lim[k][1] = p->domain_max[k] + p->bucket_side;
                }
                p->phys_min[k] = lim[k][0];
                p->phys_max[k] = lim[k][1];
                long nb = (long)((lim[k][1] - lim[k][0]) / p->bucket_side + 1);
                *nbk[k] = (int)nb;
                lim[k][1] = lim[k][0] + nb * p->bucket_side;
            }
            if (*nbk[d] < 3) { err = "periodic direction shorter than one bucket"; return MPS_E_UNSUPPORTED; }
        }
        for (int k = 0; k < 3; k++) { p->domain_min[k] = lim[k][0]; p->domain_max[k] = lim[k][1]; }
        p->inv_bucket_side = 1.0 / p->bucket_side;
    }
    p->num_buckets_xy = p->num_buckets_x * p->num_buckets_y;
    p->num_buckets_xyz = p->num_buckets_xy * p->num_buckets_z;
    p->derived = 1;
    return MPS_OK;
}

}  // namespace mpsgpu

extern "C" {

size_t mps_params_size(void) { return sizeof(mps_params); }

int mps_params_defaults(mps_params* p) {
    if (!p) return MPS_E_ARG;
    mpsgpu::paramsDefaults(p);
    return MPS_OK;
}

int mps_params_from_json(const char* json_path, mps_params* p, char* grid_path, size_t grid_cap,
                         char* mesh_path, size_t mesh_cap) {
    if (!json_path || !p) return MPS_E_ARG;
    std::string g, m, err;
    int rc = mpsgpu::paramsFromJson(json_path, p, g, m, err);
    if (rc) {
        mpsgpu::setGlobalError(err);
        return rc;
    }
    if (grid_path && grid_cap) snprintf(grid_path, grid_cap, "%s", g.c_str());
    if (mesh_path && mesh_cap) snprintf(mesh_path, mesh_cap, "%s", m.c_str());
    return MPS_OK;
}

int mps_mesh_paths_from_json(const char* json_path, char* deformable_path, size_t deformable_cap, char* forced_path,
                             size_t forced_cap) {
    if (!json_path) return MPS_E_ARG;
    std::ifstream f(json_path);
    if (!f) {
        mpsgpu::setGlobalError(std::string("cannot open ") + json_path);
        return MPS_E_IO;
    }
    std::stringstream ss;
    ss << f.rdbuf();
    std::string text = ss.str(), err;
    mpsgpu::JVal root;
    if (!mpsgpu::JParser(text).parse(root, err)) {
        mpsgpu::setGlobalError("JSON parse error: " + err);
        return MPS_E_IO;
    }
    const mpsgpu::JVal* pn = root.get("pathNames");
    if (deformable_path && deformable_cap)
        snprintf(deformable_path, deformable_cap, "%s", mpsgpu::jstr(pn, "mesh_deformable_file", "oops").c_str());
    if (forced_path && forced_cap)
        snprintf(forced_path, forced_cap, "%s", mpsgpu::jstr(pn, "mesh_forced_file", "oops").c_str());
    return MPS_OK;
}

int mps_output_from_json(const char* json_path, char* vtu_folder, size_t folder_cap, int* vtu_flags) {
    if (!json_path) return MPS_E_ARG;
    std::ifstream f(json_path);
    if (!f) { mpsgpu::setGlobalError(std::string("cannot open ") + json_path); return MPS_E_IO; }
    std::stringstream ss;
    ss << f.rdbuf();
    const std::string text = ss.str();
    mpsgpu::JVal root;
    std::string err;
    if (!mpsgpu::JParser(text).parse(root, err)) { mpsgpu::setGlobalError("JSON parse error: " + err); return MPS_E_IO; }
    const mpsgpu::JVal* fl = root.get("flags");
    const mpsgpu::JVal* vtu = fl ? fl->get("output_VTU") : nullptr;
    const mpsgpu::JVal* pn = root.get("pathNames");
    if (vtu_folder && folder_cap) snprintf(vtu_folder, folder_cap, "%s", mpsgpu::jstr(pn, "vtu_output_folder", "oops").c_str());
    if (vtu_flags) {
        int v = 0;
        if (mpsgpu::jint(vtu, "pnd", 0)) v |= MPS_VTU_PND;
        if (mpsgpu::jint(vtu, "neigh", 0)) v |= MPS_VTU_NEIGH;
        if (mpsgpu::jint(vtu, "deviation", 0)) v |= MPS_VTU_DEVIATION;
        if (mpsgpu::jint(vtu, "concentration", 0)) v |= MPS_VTU_CONCENTRATION;
        if (mpsgpu::jint(vtu, "auxiliar", 0)) v |= MPS_VTU_AUXILIAR;
        if (mpsgpu::jint(vtu, "non_newtonian", 0)) v |= MPS_VTU_NON_NEWTONIAN;
        if (mpsgpu::jint(vtu, "only_freeSurface", 0)) v |= MPS_VTU_ONLY_FREE_SURFACE;
        if (mpsgpu::jint(vtu, "type", 0)) v |= MPS_VTU_BINARY;
        *vtu_flags = v;
    }
    return MPS_OK;
}

int mps_params_derive(mps_params* p, int n, const double* pos, const int* type) {
    (void)type;
    if (!p) return MPS_E_ARG;
    std::string err;
    int rc = mpsgpu::paramsDerive(p, n, pos, err);
    if (rc) mpsgpu::setGlobalError(err);
    return rc;
}

}  // extern "C"
