// mcg_host.cpp -- host-side derivation of the device scene (FP64).
//
// Restates what the reference's constructors compute before a solve:
//   Material::Material (material.cpp:82-162): tau, dE/dT, the three pdfs,
//     their sums and the bulk conductivity; here also the sampling tables.
//   Boundary / Shape ctors (boundary.cpp:48-56, 123-187, 264-301, 391-393):
//     inward plane normal from the face's edge vectors, offset, rotation and
//     reflection matrices.
//   Parallelepiped / TriangularPrism / Tetrahedron / Prism / Pyramid faces
//     (subdomain.h:149-154, 228-233, 305-309, 398-409, 537-546), their
//     volumes (subdomain.cpp:379-441) and Subdomain::Subdomain
//     (subdomain.cpp:41-71).
//   makePair (boundary.cpp:361-369, 524-550), emitter registry
//     (subdomain.cpp:205-212, domain.cpp:88-102), Field::init column layout
//     (field.cpp:25-45) and cell volumes (subdomain.cpp:269-349).
//   FieldProblem::FieldProblem (problem.cpp:315-342).
//   postProc and normalisation (problem.cpp:439-444, 478-648).
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstring>

#include "mcg_disk.h"
#include "mcg_scene.h"

namespace mcg {

namespace {

constexpr double kPi = 3.141592653589793;       // constants.h:17
constexpr double kHbar = 1.054560652927e-034;   // constants.h:18
constexpr double kKb = 1.380648e-023;           // constants.h:19

struct Vec {
    double v[3];
    double operator[](int i) const { return v[i]; }
};

Vec mk(double a, double b, double c) { return Vec{{a, b, c}}; }
Vec plus(const Vec& a, const Vec& b) { return mk(a[0] + b[0], a[1] + b[1], a[2] + b[2]); }
Vec minus(const Vec& a, const Vec& b) { return mk(a[0] - b[0], a[1] - b[1], a[2] - b[2]); }
Vec times(double s, const Vec& a) { return mk(s * a[0], s * a[1], s * a[2]); }
double dotp(const Vec& a, const Vec& b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
Vec crossp(const Vec& a, const Vec& b) {
    return mk(a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]);
}
double len(const Vec& a) { return std::sqrt(dotp(a, a)); }
Vec unit(const Vec& a) {
    double q = dotp(a, a);
    if (q > 0.0) {
        double l = std::sqrt(q);
        return mk(a[0] / l, a[1] / l, a[2] / l);
    }
    return a;
}
// column-major 3x3 helpers
Vec column(const double* m, int c) { return mk(m[3 * c], m[3 * c + 1], m[3 * c + 2]); }
Vec matvec(const double* m, const Vec& x) {
    return mk(m[0] * x[0] + m[3] * x[1] + m[6] * x[2], m[1] * x[0] + m[4] * x[1] + m[7] * x[2],
              m[2] * x[0] + m[5] * x[1] + m[8] * x[2]);
}
double determinant(const double* m) {
    return m[0] * (m[4] * m[8] - m[7] * m[5]) + m[3] * (m[7] * m[2] - m[1] * m[8]) +
           m[6] * (m[1] * m[5] - m[4] * m[2]);
}
void invert(const double* m, double* r) {
    double d = determinant(m);
    r[0] = (m[4] * m[8] - m[7] * m[5]) / d;
    r[3] = (m[6] * m[5] - m[3] * m[8]) / d;
    r[6] = (m[3] * m[7] - m[6] * m[4]) / d;
    r[1] = (m[7] * m[2] - m[1] * m[8]) / d;
    r[4] = (m[0] * m[8] - m[6] * m[2]) / d;
    r[7] = (m[6] * m[1] - m[0] * m[7]) / d;
    r[2] = (m[1] * m[5] - m[4] * m[2]) / d;
    r[5] = (m[3] * m[2] - m[0] * m[5]) / d;
    r[8] = (m[0] * m[4] - m[3] * m[1]) / d;
}
double fro(const double* m) {
    double s = 0.0;
    for (int i = 0; i < 9; ++i) s += m[i] * m[i];
    return std::sqrt(s);
}
bool approx_mat(const double* a, const double* b) {
    double d[9];
    for (int i = 0; i < 9; ++i) d[i] = a[i] - b[i];
    return fro(d) <= 1e-12 * std::min(fro(a), fro(b));
}
bool approx_vec(const Vec& a, const Vec& b) {
    return len(minus(a, b)) <= 1e-12 * std::min(len(a), len(b));
}
bool approx_scalar(double a, double b) {   // constants.h:21-29
    double rhs = std::max(DBL_MIN, 100 * DBL_EPSILON * std::max(std::fabs(a), std::fabs(b)));
    return std::fabs(a - b) <= rhs;
}

// boundary.cpp:28-31: I - 2 n n^T
void reflection(const Vec& n, double* r) {
    for (int c = 0; c < 3; ++c)
        for (int q = 0; q < 3; ++q) r[3 * c + q] = (q == c ? 1.0 : 0.0) - 2.0 * (n[q] * n[c]);
}

// boundary.cpp:33-37: the rotation taking z onto n, built from the
// half-angle quaternion; an n opposite to z rotates about x.
void rotation_to(const Vec& n_in, double* r) {
    Vec n = unit(n_in);
    double c = n[2], qw, qx, qy, qz;
    if (c < -1.0 + 1e-12) {
        c = std::max(c, -1.0);
        double w2 = (1.0 + c) * 0.5;
        qw = std::sqrt(w2);
        qx = std::sqrt(1.0 - w2);
        qy = qz = 0.0;
    } else {
        double s = std::sqrt((1.0 + c) * 2.0);
        double invs = 1.0 / s;
        qx = -n[1] * invs;
        qy = n[0] * invs;
        qz = 0.0 * invs;
        qw = s * 0.5;
    }
    double tx = 2.0 * qx, ty = 2.0 * qy, tz = 2.0 * qz;
    double twx = tx * qw, twy = ty * qw, twz = tz * qw;
    double txx = tx * qx, txy = ty * qx, txz = tz * qx;
    double tyy = ty * qy, tyz = tz * qy, tzz = tz * qz;
    r[0] = 1.0 - (tyy + tzz);
    r[3] = txy - twz;
    r[6] = txz + twy;
    r[1] = txy + twz;
    r[4] = 1.0 - (txx + tzz);
    r[7] = tyz - twx;
    r[2] = txz - twy;
    r[5] = tyz + twx;
    r[8] = 1.0 - (txx + tyy);
}

uint32_t prob_to_thr(double p) { return p >= 1.0 ? 0xFFFFFFFFu : uint32_t(p * 4294967296.0); }

// A partially specular wall draws specular when the event's word is below
// p 2^32; its end points are the reference's own kinds, SpecBoundary at
// p = 1 and DiffBoundary at p = 0 (boundary.cpp:283-287, 308-312), so a wall
// there takes that kind and no draw can leave it.
int partial_kind(double p) { return p >= 1.0 ? MCG_SPEC : (p <= 0.0 ? MCG_DIFF : MCG_PARTIAL); }

struct FaceDef {
    Vec o, i, j;
    int shape;
    int nfan = 0;                  // SHAPE_POLY: vertices after the first,
    Vec fan[MCG_MAX_SIDES - 1];    // relative to o (Polygon<N>, boundary.cpp:197-216)
};

// Sides of a prism's or pyramid's N-gon base for a face count, 0 if out of range.
int base_sides(int cell, int nface) {
    const int n = cell == MCG_PRISM ? nface - 2 : cell == MCG_PYRAMID ? nface - 1 : 0;
    return n >= 4 && n <= MCG_MAX_SIDES ? n : 0;
}

FaceDef polygon(const Vec& o, const Vec* v, int nv) {
    FaceDef f{o, v[0], v[1], SHAPE_POLY};
    f.nfan = nv;
    for (int k = 0; k < nv; ++k) f.fan[k] = v[k];
    return f;
}

int faces_of(int cell, int nf_in, const Vec& o, const double* m, FaceDef* f, int& nface) {
    Vec a = column(m, 0), b = column(m, 1), c = column(m, 2);
    switch (cell) {
        case MCG_PARALLELEPIPED:   // Bac Lef Bot Fro Rig Top
            f[0] = {o, b, c, SHAPE_PAR};
            f[1] = {o, c, a, SHAPE_PAR};
            f[2] = {o, a, b, SHAPE_PAR};
            f[3] = {plus(o, a), c, b, SHAPE_PAR};
            f[4] = {plus(o, b), a, c, SHAPE_PAR};
            f[5] = {plus(o, c), b, a, SHAPE_PAR};
            nface = 6;
            return MCG_OK;
        case MCG_TRIPRISM:         // Bac Lef Bot Dia Top
            f[0] = {o, b, c, SHAPE_PAR};
            f[1] = {o, c, a, SHAPE_PAR};
            f[2] = {o, a, b, SHAPE_TRI};
            f[3] = {plus(o, a), c, minus(b, a), SHAPE_PAR};
            f[4] = {plus(o, c), b, a, SHAPE_TRI};
            nface = 5;
            return MCG_OK;
        case MCG_TETRAHEDRON:      // Bac Lef Bot Dia
            f[0] = {o, b, c, SHAPE_TRI};
            f[1] = {o, c, a, SHAPE_TRI};
            f[2] = {o, a, b, SHAPE_TRI};
            f[3] = {plus(o, a), minus(c, a), minus(b, a), SHAPE_TRI};
            nface = 4;
            return MCG_OK;
        case MCG_PRISM:            // Bot Top Sid0..Sid(N-1) (subdomain.h:398-409, 462-479)
        case MCG_PYRAMID: {        // Bot Sid0..Sid(N-1) (subdomain.h:537-546, 598-617)
            const int n = base_sides(cell, nf_in);
            if (!n) return set_error(MCG_EINVAL, "a prism or pyramid base needs 4 to 9 sides");
            Vec v[MCG_MAX_SIDES];
            for (int k = 0; k < n; ++k) v[k] = column(m, k);
            Vec rev[MCG_MAX_SIDES - 1];
            for (int k = 0; k < n - 1; ++k) rev[k] = v[n - 1 - k];
            int q = 0;
            f[q++] = polygon(o, v + 1, n - 1);                          // base, vertices 1..N-1
            if (cell == MCG_PRISM) f[q++] = polygon(plus(o, a), rev, n - 1);   // top, reversed
            for (int s = 0; s < n; ++s) {
                // side s runs from vertex s to vertex s + 1 (vertex N is the origin's)
                const Vec p = s == 0 ? mk(0.0, 0.0, 0.0) : v[s];
                const Vec j = minus(s != n - 1 ? v[s + 1] : mk(0.0, 0.0, 0.0), p);
                f[q++] = cell == MCG_PRISM ? FaceDef{plus(o, p), a, j, SHAPE_PAR}
                                           : FaceDef{plus(o, p), minus(a, p), j, SHAPE_TRI};
            }
            nface = q;
            return MCG_OK;
        }
    }
    return set_error(MCG_EINVAL, "unknown cell kind");
}

// Volume pieces of a prism (div 2) or pyramid (div 6) as the reference's
// PrismImpl / PyramidImpl::volume compute them (subdomain.cpp:385-394,
// 417-426): term i is (m_i x m_(i+1)) . m_0 / div for i = 0..N-3, so term 0
// is zero (m_0 is the axis) and the fan's last triangle is left out; their
// drawPos picks term i and draws in the piece with edges (m_(i+1), m_(i+2),
// m_0) (subdomain.cpp:401-409, 433-441).  Both are kept as written.
void volume_terms(const double* m, int n, double div, double* w) {
    const Vec a = column(m, 0);
    for (int i = 0; i < n - 2; ++i) w[i] = dotp(crossp(column(m, i), column(m, i + 1)), a) / div;
}

// Piece table of an emitter: the pieces' edges and their cumulative,
// normalised weights.  A weight that is negative only by rounding (the
// first volume term, (m_0 x m_1) . m_0) counts as zero; pcum is exactly 1
// from the last piece of nonzero weight on, so no later piece is drawn.
int set_pieces(DevEmitter<double>& e, const Vec* pa, const Vec* pb, const double* w, int n) {
    double total = 0.0, mag = 0.0;
    for (int k = 0; k < n; ++k) mag += std::fabs(w[k]);
    int last = -1;
    for (int k = 0; k < n; ++k) {
        if (w[k] < -1e-12 * mag) return set_error(MCG_EINVAL, "Volume too small, check vector order");
        if (w[k] > 0.0) {
            total += w[k];
            last = k;
        }
    }
    if (last < 0) return set_error(MCG_EINVAL, "Volume too small, check vector order");
    double run = 0.0;
    e.npiece = n;
    for (int k = 0; k < n; ++k) {
        run += std::max(w[k], 0.0);
        e.pcum[k] = k >= last ? 1.0 : run / total;
        for (int r = 0; r < 3; ++r) {
            e.pa[k][r] = pa[k][r];
            e.pb[k][r] = pb[k][r];
        }
    }
    return MCG_OK;
}

double volume_of_cell(int cell, double vol, const int64_t shape[3], const int64_t idx[3]) {
    if (cell == MCG_PARALLELEPIPED) return vol / double(shape[0] * shape[1] * shape[2]);
    if (cell == MCG_PRISM || cell == MCG_PYRAMID) return vol;   // subdomain.cpp:396-399, 428-431
    if (cell == MCG_TRIPRISM) {
        double s0 = double(shape[0]), s1 = double(shape[1]);
        double f0 = 1.0 - (double(idx[0]) / s0 + double(idx[1]) / s1);
        if (f0 <= 0.0) return 0.0;
        if (f0 - (1.0 / s0 + 1.0 / s1) >= 0.0) return vol / double(shape[0] * shape[1] * shape[2]);
        double frac = f0 * f0;
        double fa = f0 - (1.0 / s0 + 0.0 / s1);
        if (fa > 0.0) frac += -1.0 * (fa * fa);
        double fb = f0 - (0.0 / s0 + 1.0 / s1);
        if (fb > 0.0) frac += -1.0 * (fb * fb);
        return vol * frac / (2.0 * double(shape[2]));
    }
    double s0 = double(shape[0]), s1 = double(shape[1]), s2 = double(shape[2]);
    double f0 = 1.0 - (double(idx[0]) / s0 + double(idx[1]) / s1 + double(idx[2]) / s2);
    if (f0 <= 0.0) return 0.0;
    if (f0 - (1.0 / s0 + 1.0 / s1 + 1.0 / s2) >= 0.0)
        return vol / double(shape[0] * shape[1] * shape[2]);
    static const int corner[6][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1},
                                     {0, 1, 1}, {1, 0, 1}, {1, 1, 0}};
    double frac = f0 * f0 * f0;
    for (int k = 0; k < 6; ++k) {
        double sg = ((corner[k][0] + corner[k][1] + corner[k][2]) % 2 == 0) ? 1.0 : -1.0;
        double f = f0 - (corner[k][0] / s0 + corner[k][1] / s1 + corner[k][2] / s2);
        if (f > 0.0) frac += sg * (f * f * f);
    }
    return vol * frac / 6.0;
}

}  // namespace

int build_alias(const double* pdf, int64_t n, std::vector<uint64_t>& table) {
    // Walker's alias method, tables built with Vose's small/large worklists.
    if (n <= 0 || n > 0xFFFFFFFFll) return set_error(MCG_EINVAL, "distribution size out of range");
    double total = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        if (!(pdf[i] >= 0.0) || !std::isfinite(pdf[i]))
            return set_error(MCG_EINVAL, "distribution has a negative or non-finite weight");
        total += pdf[i];
    }
    if (!(total > 0.0)) return set_error(MCG_EINVAL, "distribution has zero total weight");
    std::vector<double> q(n), prob(n, 1.0);
    std::vector<uint32_t> al(n);
    std::vector<int64_t> lo, hi;
    lo.reserve(n);
    hi.reserve(n);
    for (int64_t i = 0; i < n; ++i) {
        al[i] = uint32_t(i);
        q[i] = pdf[i] * double(n) / total;
        (q[i] < 1.0 ? lo : hi).push_back(i);
    }
    while (!lo.empty() && !hi.empty()) {
        int64_t s = lo.back();
        lo.pop_back();
        int64_t l = hi.back();
        hi.pop_back();
        prob[s] = q[s];
        al[s] = uint32_t(l);
        q[l] = (q[l] + q[s]) - 1.0;
        (q[l] < 1.0 ? lo : hi).push_back(l);
    }
    table.resize(n);
    for (int64_t i = 0; i < n; ++i)
        table[i] = uint64_t(prob_to_thr(prob[i])) | (uint64_t(al[i]) << 32);
    return MCG_OK;
}

int build_material(const mcg_material_spec* spec, HostScene& s) {
    if (!spec || !spec->disp || !spec->relax) return set_error(MCG_EINVAL, "material spec is NULL");
    if (spec->nw <= 0 || spec->np <= 0) return set_error(MCG_EINVAL, "Invalid dispersion file");
    if (!(spec->temp > 0.0)) return set_error(MCG_EINVAL, "temperature must be positive");
    const int64_t nw = spec->nw, np = spec->np, ncol = 2 + 2 * np, n = nw * np;
    const double T = spec->temp;
    s.nw = nw;
    s.np = np;
    s.temp = T;
    std::vector<double> omega(nw), dw(nw), dos(n);
    s.vel.assign(n, 0.0);
    for (int64_t w = 0; w < nw; ++w) {
        const double* r = spec->disp + w * ncol;
        omega[w] = r[0];
        dw[w] = r[1];
        for (int64_t p = 0; p < np; ++p) {
            s.vel[w * np + p] = r[2 + 2 * p];
            dos[w * np + p] = r[3 + 2 * p];
        }
    }
    // relaxation times: 1/tau = sum over mechanisms (material.cpp:116-134)
    s.tau.assign(n, 0.0);
    std::vector<double> rate(nw);
    // omega^c1 once per distinct exponent (branches and mechanisms share
    // them): the same std::pow values, computed fewer times
    std::vector<std::pair<double, std::vector<double>>> powers;
    auto omega_pow = [&](double e) -> const std::vector<double>& {
        for (const auto& q : powers)
            if (q.first == e) return q.second;
        std::vector<double> v(nw);
        for (int64_t w = 0; w < nw; ++w) v[w] = std::pow(omega[w], e);
        powers.emplace_back(e, std::move(v));
        return powers.back().second;
    };
    for (int64_t p = 0; p < np; ++p) {
        std::fill(rate.begin(), rate.end(), 0.0);
        for (int mech = 0; mech < 2; ++mech) {
            const double* c = spec->relax + p * 8 + 4 * mech;
            if (c[0] < 0.0) return set_error(MCG_EINVAL, "Scattering times cannot be negative");
            if (c[0] <= DBL_MIN) continue;
            const double tp = std::pow(T, c[2]), ex = std::exp(-c[3] / T);
            const std::vector<double>& pw = omega_pow(c[1]);
            for (int64_t w = 0; w < nw; ++w)
                rate[w] = rate[w] + ((c[0] * pw[w]) * tp) * ex;
        }
        for (int64_t w = 0; w < nw; ++w) {
            double t = 1.0 / rate[w];
            if (!std::isfinite(t))
                return set_error(MCG_EINVAL, "Scattering time model produced infinite values");
            s.tau[w * np + p] = t;
        }
    }
    // modal heat capacity k_B (x / 2 sinh(x/2))^2 (material.cpp:136-145)
    const double xscale = kHbar / (kKb * T);
    s.energy.assign(n, 0.0);
    s.flux.assign(n, 0.0);
    s.scat.assign(n, 0.0);
    s.energy_sum = s.flux_sum = s.scat_sum = 0.0;
    double ksum = 0.0;
    for (int64_t w = 0; w < nw; ++w) {
        double x = xscale * omega[w];
        double cv;
        if (std::fabs(x) < DBL_EPSILON) {
            cv = 1.0 - x * x / 12.0;
        } else {
            double h = x / (2.0 * std::sinh(x / 2.0));
            cv = h * h;
        }
        cv = kKb * cv;
        for (int64_t p = 0; p < np; ++p) {
            const int64_t i = w * np + p;
            const double e = (cv * dos[i]) * dw[w];
            if (e < 0.0) return set_error(MCG_EINVAL, "negative density of states");
            s.energy[i] = e;
            s.flux[i] = s.vel[i] * e;
            s.scat[i] = e / s.tau[i];
            s.energy_sum += e;
            s.flux_sum += s.flux[i];
            s.scat_sum += s.scat[i];
            ksum += (s.tau[i] * (s.vel[i] * s.vel[i])) * e;
        }
    }
    s.cond = ksum / 3.0;
    // draw support and per-mode flight constants
    s.mode_cell.clear();
    s.mode_vel.clear();
    s.mode_mfp.clear();
    std::vector<double> fpdf, spdf;
    for (int64_t i = 0; i < n; ++i) {
        if (!(s.energy[i] > 0.0)) continue;
        if ((s.flux[i] > 0.0 || s.scat[i] > 0.0) && !(s.vel[i] > DBL_MIN))
            return set_error(MCG_EINVAL, "Velocity must be nonzero");
        s.mode_cell.push_back(i);
        s.mode_vel.push_back(s.vel[i]);
        s.mode_mfp.push_back(s.vel[i] * s.tau[i]);
        fpdf.push_back(s.flux[i]);
        spdf.push_back(s.scat[i]);
    }
    if (s.mode_cell.empty()) return set_error(MCG_EINVAL, "material has no populated modes");
    int rc = build_alias(fpdf.data(), int64_t(fpdf.size()), s.flux_alias);
    if (rc) return rc;
    return build_alias(spdf.data(), int64_t(spdf.size()), s.scat_alias);
}

int build_domain(const mcg_domain_spec* spec, HostScene& s) {
    if (!spec || spec->nsdom <= 0 || !spec->sdom)
        return set_error(MCG_EINVAL, "Domain setup not complete");
    s.planes.clear();
    s.sdoms.clear();
    s.emitters.clear();
    s.emit_weight.clear();
    s.inter_list.clear();
    s.pores.clear();
    s.has_1d = s.has_grid = false;
    s.grid_2d = true;
    s.grid_cells = 0;
    const int ns = spec->nsdom;
    std::vector<double> face_temp;            // pre-pairing temperature per plane
    std::vector<double> face_area;
    std::vector<FaceDef> face_geo;
    std::vector<double> sdom_vol(ns);
    std::vector<std::vector<int64_t>> shapes(ns, std::vector<int64_t>(3));
    std::vector<std::vector<int32_t>> inter_pairs;
    std::vector<int32_t> peri_pair;
    for (int si = 0; si < ns; ++si) {
        const mcg_subdomain_spec& c = spec->sdom[si];
        if (c.cell < MCG_PARALLELEPIPED || c.cell > MCG_PYRAMID)
            return set_error(MCG_EINVAL, "unknown cell kind");
        const bool ngon = c.cell == MCG_PRISM || c.cell == MCG_PYRAMID;
        const int nside = ngon ? base_sides(c.cell, c.nface) : 3;
        if (ngon && !nside) return set_error(MCG_EINVAL, "a prism or pyramid base needs 4 to 9 sides");
        const int want = c.cell == MCG_PARALLELEPIPED ? 6 : c.cell == MCG_TRIPRISM ? 5
                       : c.cell == MCG_TETRAHEDRON ? 4 : c.nface;
        if (c.nface != want) return set_error(MCG_EINVAL, "Incorrect number of temperatures");
        // the matrix the cell's coordinates use: the edges, or for an N-gon
        // cell (m_1, m_(N-1), m_0) (subdomain.cpp:379-383, 411-415)
        double base[9];
        if (ngon) {
            const int cols[3] = {1, nside - 1, 0};
            for (int k = 0; k < 3; ++k)
                for (int r = 0; r < 3; ++r) base[3 * k + r] = c.mat[3 * cols[k] + r];
        } else {
            std::memcpy(base, c.mat, sizeof(base));
        }
        double vol;
        if (ngon) {
            double w[MCG_MAX_SIDES];
            volume_terms(c.mat, nside, c.cell == MCG_PRISM ? 2.0 : 6.0, w);
            vol = 0.0;
            for (int k = 0; k < nside - 2; ++k) vol += w[k];
        } else {
            const double det = determinant(c.mat);
            vol = c.cell == MCG_PARALLELEPIPED ? det : c.cell == MCG_TRIPRISM ? det / 2.0 : det / 6.0;
        }
        if (!(vol >= DBL_MIN)) return set_error(MCG_EINVAL, "Volume too small, check vector order");
        sdom_vol[si] = vol;
        DevSdom<double> d{};
        const Vec o = mk(c.origin[0], c.origin[1], c.origin[2]);
        for (int k = 0; k < 3; ++k) d.o[k] = o[k];
        invert(base, d.inv);
        // accumulation flag and grid shape (subdomain.cpp:41-71); an N-gon
        // cell passes (-1, -1, -1) or (0, 0, 0) (subdomain.h:388-391)
        int64_t div[3];
        for (int k = 0; k < 3; ++k) div[k] = ngon ? (c.div[0] < 0 ? -1 : 0) : c.div[k];
        int positive = 0;
        bool negative = false;
        for (int k = 0; k < 3; ++k) {
            negative |= div[k] < 0;
            positive += div[k] > 0;
            shapes[si][k] = std::max<int64_t>(div[k], 1);
            d.divf[k] = double(div[k]);
        }
        if (negative) {
            d.accum = -2;
            for (int k = 0; k < 3; ++k) shapes[si][k] = 0;
        } else if (positive == 0) {
            d.accum = -1;
        } else if (positive == 1) {
            int axis = 0;
            for (int k = 1; k < 3; ++k)
                if (div[k] > div[axis]) axis = k;
            d.accum = axis;
            s.has_1d = true;
        } else {
            d.accum = positive + 1;
            s.has_grid = true;
            s.grid_2d = s.grid_2d && positive == 2;
            double cells = 1.0;
            for (int k = 0; k < 3; ++k) cells *= double(std::max<int64_t>(div[k], 1));
            s.grid_cells = std::max<int64_t>(s.grid_cells, int64_t(std::min(cells, 1e18)));
            int g = 0;
            for (int k = 0; k < 3; ++k)
                if (div[k] > 0 && g < 2) d.gax[g++] = k;
        }
        for (int k = 0; k < 3; ++k) {
            if (shapes[si][k] > (1ll << 30)) return set_error(MCG_ERANGE, "too many tally divisions");
            d.max[k] = int32_t(negative ? 0 : shapes[si][k] - 1);
        }
        double colmax = 0.0;
        for (int k = 0; k < 3; ++k) colmax = std::max(colmax, len(column(base, k)));
        d.eps = 100.0 * DBL_EPSILON * colmax;
        d.pore = -1;
        FaceDef fd[MCG_MAX_FACES];
        int nf = 0;
        int rc = faces_of(c.cell, c.nface, o, c.mat, fd, nf);
        if (rc) return rc;
        d.pbeg = int32_t(s.planes.size());
        for (int f = 0; f < nf; ++f) {
            const mcg_face_spec& fs = c.face[f];
            if (fs.kind < MCG_SPEC || fs.kind > MCG_PARTIAL)
                return set_error(MCG_EINVAL, "unknown boundary kind");
            const bool emits = fs.kind == MCG_ISOT || fs.kind == MCG_PERI;
            if (!emits && fs.temp != 0.0) return set_error(MCG_EINVAL, "Cannot specify temperature");
            DevPlane<double> P{};
            const Vec cr = crossp(fd[f].i, fd[f].j);
            const Vec n = unit(cr);
            double area = fd[f].shape == SHAPE_PAR ? len(cr) : len(cr) / 2.0;
            if (fd[f].shape == SHAPE_POLY) {
                // the fan's triangles, each of positive area and facing the
                // polygon's way (boundary.cpp:197-216)
                area = 0.0;
                for (int k = 0; k + 1 < fd[f].nfan; ++k) {
                    const Vec t = crossp(fd[f].fan[k], fd[f].fan[k + 1]);
                    if (!(len(t) / 2.0 > DBL_MIN)) return set_error(MCG_EINVAL, "Area too small");
                    if (!approx_vec(unit(t), n)) return set_error(MCG_EINVAL, "Normals are inconsistent");
                    area += len(t) / 2.0;
                }
            }
            if (area == 0.0) return set_error(MCG_EINVAL, "Shape not initialized");
            for (int k = 0; k < 3; ++k) P.n[k] = n[k];
            P.c = -dotp(n, fd[f].o);
            rotation_to(n, P.rot);
            reflection(n, P.refl);
            P.kind = fs.kind;
            P.sdom = si;
            P.target = si;
            P.ibeg = P.icnt = 0;
            if (fs.kind == MCG_PARTIAL) {
                if (!(fs.specularity >= 0.0 && fs.specularity <= 1.0))
                    return set_error(MCG_EINVAL, "specularity outside [0, 1]");
                P.kind = partial_kind(fs.specularity);
                P.spec_thr = prob_to_thr(fs.specularity);
            }
            s.planes.push_back(P);
            face_temp.push_back(fs.temp);
            face_area.push_back(area);
            face_geo.push_back(fd[f]);
            inter_pairs.emplace_back();
            peri_pair.push_back(-1);
        }
        d.pend = int32_t(s.planes.size());
        // six faces whose normals are coordinate axes: a box, advected axis by
        // axis in FP32 flight (a face with inward normal +e_a bounds p_a from
        // below at -c, one with -e_a from above at c)
        d.box = 0;
        if (nf == 6) {
            int lo = 0, hi = 0;
            for (int f = 0; f < 6; ++f) {
                const DevPlane<double>& P = s.planes[d.pbeg + f];
                int ax = -1, zeros = 0;
                for (int k = 0; k < 3; ++k) {
                    if (P.n[k] == 0.0) ++zeros;
                    else if (P.n[k] == 1.0 || P.n[k] == -1.0) ax = k;
                }
                if (zeros != 2 || ax < 0) break;
                if (P.n[ax] > 0.0) {
                    d.bmin[ax] = -P.c;
                    d.fmin[ax] = f;
                    lo |= 1 << ax;
                } else {
                    d.bmax[ax] = P.c;
                    d.fmax[ax] = f;
                    hi |= 1 << ax;
                }
            }
            d.box = lo == 7 && hi == 7 ? 1 : 0;
        }
        s.sdoms.push_back(d);
    }
    std::vector<double> temp_now(face_temp);
    // pairs
    for (int k = 0; k < spec->npair; ++k) {
        const mcg_pair_spec& ps = spec->pair[k];
        if (ps.sdom_a < 0 || ps.sdom_a >= ns || ps.sdom_b < 0 || ps.sdom_b >= ns)
            return set_error(MCG_EINVAL, "pair names an unknown subdomain");
        const DevSdom<double>& A = s.sdoms[ps.sdom_a];
        const DevSdom<double>& B = s.sdoms[ps.sdom_b];
        if (ps.face_a < 0 || ps.face_a >= A.pend - A.pbeg || ps.face_b < 0 ||
            ps.face_b >= B.pend - B.pbeg)
            return set_error(MCG_EINVAL, "pair names an unknown face");
        const int ia = A.pbeg + ps.face_a, ib = B.pbeg + ps.face_b;
        DevPlane<double>& a = s.planes[ia];
        DevPlane<double>& b = s.planes[ib];
        const Vec na = mk(a.n[0], a.n[1], a.n[2]), nb = mk(b.n[0], b.n[1], b.n[2]);
        if (a.kind == MCG_PERI && b.kind == MCG_PERI) {
            double R[9], Rt[9];
            bool zero = true;
            for (int i = 0; i < 9; ++i) zero = zero && ps.rot[i] == 0.0;
            for (int i = 0; i < 9; ++i) R[i] = zero ? ((i % 4 == 0) ? 1.0 : 0.0) : ps.rot[i];
            for (int c = 0; c < 3; ++c)
                for (int r = 0; r < 3; ++r) Rt[3 * c + r] = R[3 * r + c];
            if (peri_pair[ia] >= 0 || peri_pair[ib] >= 0)
                return set_error(MCG_EINVAL, "Boundary already paired");
            double RRt[9], I[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
            for (int c = 0; c < 3; ++c) {
                Vec v = matvec(R, column(Rt, c));
                for (int r = 0; r < 3; ++r) RRt[3 * c + r] = v[r];
            }
            if (!approx_mat(I, RRt)) return set_error(MCG_EINVAL, "Rotation matrix must be orthogonal");
            if (!approx_vec(matvec(R, na), times(-1.0, nb)))
                return set_error(MCG_EINVAL, "Rotation matrix incorrect");
            const Vec t = mk(ps.transl[0], ps.transl[1], ps.transl[2]);
            if (!approx_scalar(dotp(nb, plus(times(-a.c, matvec(R, na)), t)), -b.c))
                return set_error(MCG_EINVAL, "Translation vector incorrect");
            const double t1 = temp_now[ia], t2 = temp_now[ib];
            temp_now[ia] = t1 - t2;
            temp_now[ib] = t2 - t1;
            std::memcpy(a.prot, R, sizeof(R));
            std::memcpy(b.prot, Rt, sizeof(Rt));
            const Vec tb = times(-1.0, matvec(Rt, t));
            for (int i = 0; i < 3; ++i) {
                a.pt[i] = t[i];
                b.pt[i] = tb[i];
            }
            a.target = b.sdom;
            b.target = a.sdom;
            peri_pair[ia] = ib;
            peri_pair[ib] = ia;
        } else if (a.kind == MCG_INTER && b.kind == MCG_INTER) {
            if (!approx_vec(na, times(-1.0, nb)))
                return set_error(MCG_EINVAL, "Boundary normals not antiparallel");
            if (!approx_scalar(a.c, -b.c)) return set_error(MCG_EINVAL, "Boundaries planes not the same");
            inter_pairs[ia].push_back(b.sdom);
            inter_pairs[ib].push_back(a.sdom);
        } else {
            return set_error(MCG_EINVAL, "pair joins faces that are not both Peri or both Inter");
        }
    }
    for (size_t i = 0; i < s.planes.size(); ++i) {
        DevPlane<double>& P = s.planes[i];
        if (P.kind == MCG_PERI && peri_pair[i] < 0) return set_error(MCG_EINVAL, "Domain setup not complete");
        if (P.kind == MCG_INTER) {
            if (inter_pairs[i].empty()) return set_error(MCG_EINVAL, "Domain setup not complete");
            P.ibeg = int32_t(s.inter_list.size());
            P.icnt = int32_t(inter_pairs[i].size());
            s.inter_list.insert(s.inter_list.end(), inter_pairs[i].begin(), inter_pairs[i].end());
            // a face with one partner carries it in the plane record, so a
            // crossing reads no partner list
            if (P.icnt == 1) P.target = inter_pairs[i].front();
        }
    }
    // pores (cylinder along z, cut out of one subdomain)
    for (int k = 0; k < spec->npore; ++k) {
        const mcg_pore_spec& ps = spec->pore[k];
        if (ps.sdom < 0 || ps.sdom >= ns) return set_error(MCG_EINVAL, "pore names an unknown subdomain");
        if (!(ps.radius > 0.0)) return set_error(MCG_EINVAL, "pore radius must be positive");
        if (ps.kind != MCG_DIFF && ps.kind != MCG_SPEC && ps.kind != MCG_PARTIAL)
            return set_error(MCG_EINVAL, "pore wall must be Diff, Spec or Partial");
        if (s.sdoms[ps.sdom].pore >= 0) return set_error(MCG_EINVAL, "one pore per subdomain");
        const mcg_subdomain_spec& c = spec->sdom[ps.sdom];
        const double* m = c.mat;
        if (c.cell != MCG_PARALLELEPIPED || m[1] != 0.0 || m[2] != 0.0 || m[3] != 0.0 || m[5] != 0.0 ||
            m[6] != 0.0 || m[7] != 0.0)
            return set_error(MCG_EINVAL, "a pore needs an axis-aligned box cell");
        if (ps.center[0] - ps.radius < c.origin[0] || ps.center[0] + ps.radius > c.origin[0] + m[0] ||
            ps.center[1] - ps.radius < c.origin[1] || ps.center[1] + ps.radius > c.origin[1] + m[4])
            return set_error(MCG_EINVAL, "pore must lie inside the cell");
        if (ps.kind == MCG_PARTIAL && !(ps.specularity >= 0.0 && ps.specularity <= 1.0))
            return set_error(MCG_EINVAL, "specularity outside [0, 1]");
        DevPore<double> P{};
        P.cx = ps.center[0];
        P.cy = ps.center[1];
        P.r = ps.radius;
        P.r2 = ps.radius * ps.radius;
        P.eps = s.sdoms[ps.sdom].eps;
        P.kind = ps.kind == MCG_PARTIAL ? partial_kind(ps.specularity) : ps.kind;
        P.spec_thr = ps.kind == MCG_PARTIAL ? prob_to_thr(ps.specularity)
                                            : (ps.kind == MCG_SPEC ? 0xFFFFFFFFu : 0u);
        // the pore runs through the cell's faces across its axis: a face
        // there keeps the hole (Peri back onto this cell) or closes it; an
        // Inter face, or a Peri face onto another cell, would let bundles in
        // through the hole
        for (int b = s.sdoms[ps.sdom].pbeg; b < s.sdoms[ps.sdom].pend; ++b) {
            const DevPlane<double>& F = s.planes[b];
            if (F.n[2] == 0.0) continue;
            if (F.kind == MCG_INTER) return set_error(MCG_EINVAL, "a pore cannot open onto an Inter face");
            if (F.kind == MCG_PERI && F.target != ps.sdom)
                return set_error(MCG_EINVAL, "a pore's Peri faces must pair within its cell");
        }
        s.sdoms[ps.sdom].pore = int32_t(s.pores.size());
        s.pores.push_back(P);
    }
    // emitter registry: the subdomain first, then its faces in order; a face
    // is registered by its temperature before pairing (subdomain.cpp:205-212)
    for (int si = 0; si < ns; ++si) {
        const mcg_subdomain_spec& c = spec->sdom[si];
        const Vec g = mk(c.grad_t[0], c.grad_t[1], c.grad_t[2]);
        // a pore removes its cylinder from the volume that emits
        const int pk = s.sdoms[si].pore;
        const double vmat = pk < 0 ? sdom_vol[si]
                                   : sdom_vol[si] - 3.141592653589793 * (s.pores[pk].r * s.pores[pk].r) * c.mat[8];
        const double wv = 2.0 * vmat * len(g);
        if (wv != 0.0) {
            DevEmitter<double> e{};
            for (int k = 0; k < 3; ++k) {
                e.o[k] = c.origin[k];
                e.grad[k] = g[k];
            }
            std::memcpy(e.mat, c.mat, sizeof(e.mat));
            rotation_to(unit(g), e.rot);
            e.volume = 1;
            e.shape = c.cell;
            e.sdom = si;
            e.sign = 0;
            if (c.cell == MCG_PRISM || c.cell == MCG_PYRAMID) {
                // piece i: edges (m_(i+1), m_(i+2), m_0), weight volume term i
                const int n = base_sides(c.cell, c.nface);
                double w[MCG_MAX_SIDES];
                volume_terms(c.mat, n, c.cell == MCG_PRISM ? 2.0 : 6.0, w);
                Vec pa[kMaxPieces], pb[kMaxPieces];
                for (int k = 0; k < n - 2; ++k) {
                    pa[k] = column(c.mat, k + 1);
                    pb[k] = column(c.mat, k + 2);
                }
                const Vec a = column(c.mat, 0);
                for (int r = 0; r < 3; ++r) e.mat[6 + r] = a[r];
                const int rc = set_pieces(e, pa, pb, w, n - 2);
                if (rc) return rc;
            }
            s.emitters.push_back(e);
            s.emit_weight.push_back(wv);
        }
        const DevSdom<double>& d = s.sdoms[si];
        for (int pi = d.pbeg; pi < d.pend; ++pi) {
            const DevPlane<double>& P = s.planes[pi];
            if (!(P.kind == MCG_ISOT || P.kind == MCG_PERI)) continue;
            if (!(face_area[pi] * std::fabs(face_temp[pi]) != 0.0)) continue;
            DevEmitter<double> e{};
            const FaceDef& f = face_geo[pi];
            for (int k = 0; k < 3; ++k) {
                e.o[k] = f.o[k];
                e.si[k] = f.i[k];
                e.sj[k] = f.j[k];
            }
            std::memcpy(e.rot, P.rot, sizeof(e.rot));
            e.volume = 0;
            e.shape = f.shape;
            e.sdom = si;
            e.sign = temp_now[pi] >= 0.0 ? 1 : 0;
            if (f.shape == SHAPE_POLY) {   // the fan's triangles by area (boundary.cpp:238-251)
                double w[kMaxPieces];
                for (int k = 0; k + 1 < f.nfan; ++k) w[k] = len(crossp(f.fan[k], f.fan[k + 1])) / 2.0;
                const int rc = set_pieces(e, f.fan, f.fan + 1, w, f.nfan - 1);
                if (rc) return rc;
            }
            s.emitters.push_back(e);
            // a face across the pore's axis has its disk cut out
            const double area = pk >= 0 && P.n[2] != 0.0
                                    ? face_area[pi] - 3.141592653589793 * (s.pores[pk].r * s.pores[pk].r)
                                    : face_area[pi];
            s.emit_weight.push_back(area * std::fabs(temp_now[pi]));
        }
    }
    // tally columns and per-column cell volumes
    s.cols = 0;
    s.cell_vol.clear();
    for (int si = 0; si < ns; ++si) {
        DevSdom<double>& d = s.sdoms[si];
        const int64_t* sh = shapes[si].data();
        const int64_t prod = sh[0] * sh[1] * sh[2];
        d.col_off = int32_t(s.cols);
        d.stride[0] = 1;
        d.stride[1] = int32_t(sh[0]);
        d.stride[2] = int32_t(sh[0] * sh[1]);
        if (prod == 0) continue;
        for (int64_t k = 0; k < sh[2]; ++k)
            for (int64_t j = 0; j < sh[1]; ++j)
                for (int64_t i = 0; i < sh[0]; ++i) {
                    const int64_t idx[3] = {i, j, k};
                    double v = volume_of_cell(spec->sdom[si].cell, sdom_vol[si], sh, idx);
                    if (d.pore >= 0) {   // material part of a cell the pore cuts
                        const DevPore<double>& P = s.pores[d.pore];
                        const mcg_subdomain_spec& c = spec->sdom[si];
                        const double lx = c.mat[0], ly = c.mat[4];
                        const double x0 = c.origin[0] + lx * (double(i) / double(sh[0]));
                        const double x1 = c.origin[0] + lx * (double(i + 1) / double(sh[0]));
                        const double y0 = c.origin[1] + ly * (double(j) / double(sh[1]));
                        const double y1 = c.origin[1] + ly * (double(j + 1) / double(sh[1]));
                        const double in = mcg_disk::rect_in_disk(x0, x1, y0, y1, P.cx, P.cy, P.r);
                        v = v * (1.0 - in / ((x1 - x0) * (y1 - y0)));
                    }
                    s.cell_vol.push_back(v);
                }
        s.cols += prod;
        if (s.cols > (1ll << 30)) return set_error(MCG_ERANGE, "too many tally columns");
    }
    return MCG_OK;
}

int build_problem(const mcg_problem_spec* spec, const HostScene& s, HostProblem& p, bool trace_only) {
    if (!spec) return set_error(MCG_EINVAL, "problem spec is NULL");
    if (spec->tally < MCG_TEMP || spec->tally > MCG_CUMFLUX) return set_error(MCG_EINVAL, "Invalid problem");
    if (spec->precision != MCG_FP64 && spec->precision != MCG_FP32_FLIGHT)
        return set_error(MCG_EINVAL, "unknown precision");
    if (spec->nemit <= 0) return set_error(MCG_EINVAL, "nemit must be positive");
    if (spec->maxscat <= 0) return set_error(MCG_EINVAL, "maxscat must be positive");
    if (spec->maxloop < 0) return set_error(MCG_EINVAL, "maxloop must be non-negative");
    // maxloop_ = loopFactor_ * maxscat_ (problem.h:124) must stay an int64
    if (spec->maxloop == 0 && spec->maxscat > INT64_MAX / 100)
        return set_error(MCG_ERANGE, "maxscat too large for the default maxloop (100 maxscat)");
    if (s.emitters.empty() && !trace_only) return set_error(MCG_EINVAL, "No phonons to emit");
    p.tally = spec->tally;
    p.precision = spec->precision;
    p.size = spec->size;
    p.step = 0;
    switch (spec->tally) {
        case MCG_TEMP: p.rows = 1; break;
        case MCG_FLUX: p.rows = 3; break;
        case MCG_MULTI: p.rows = 4; break;
        case MCG_CUMTEMP: p.rows = spec->size + 1; break;
        default: p.rows = 3 * (spec->size + 1); break;
    }
    if (spec->tally == MCG_CUMTEMP || spec->tally == MCG_CUMFLUX) {
        if (spec->size <= 0) return set_error(MCG_EINVAL, "size must be positive");
        p.step = (spec->maxscat - 1) / spec->size + (((spec->maxscat - 1) % spec->size) != 0 ? 1 : 0);
        if (p.step <= 0) return set_error(MCG_EINVAL, "maxscat must exceed 1 for a cumulative tally");
    }
    p.maxscat = spec->maxscat;
    p.maxloop = spec->maxloop != 0 ? spec->maxloop : 100 * spec->maxscat;
    p.seed = spec->seed;
    if (trace_only) {
        p.weight_sum = p.power = 0.0;
        p.emit_count.clear();
        p.emit_cdf.clear();
        p.nemit = spec->nemit;
        p.first = spec->first_bundle;
        p.count = 1;
        return MCG_OK;
    }
    double wsum = 0.0;
    for (double w : s.emit_weight) {
        if (!(w > 0.0)) return set_error(MCG_EINVAL, "No phonons to emit");
        wsum += w;
    }
    p.weight_sum = wsum;
    p.emit_count.clear();
    p.emit_cdf.clear();
    p.nemit = 0;
    for (double w : s.emit_weight) {
        const double r = std::ceil(w / wsum * double(spec->nemit) - 0.5);
        const int64_t c = std::max<int64_t>(1, int64_t(r));
        p.emit_count.push_back(c);
        p.nemit += c;
        p.emit_cdf.push_back(p.nemit);
    }
    p.power = wsum / double(p.nemit) * s.flux_sum / 4.0;
    p.first = spec->first_bundle;
    p.count = spec->bundle_count == 0 ? p.nemit - p.first : spec->bundle_count;
    if (p.first < 0 || p.count < 0 || p.first + p.count > p.nemit)
        return set_error(MCG_ERANGE, "bundle range outside [0, nemit)");
    return MCG_OK;
}

// direct / diff are rows x cols and rows x (cols + 1), row-major raw sums;
// out is rows x cols column-major, the reference's solve() value.
void finalise_field(const HostScene& s, const HostProblem& p, const double* direct,
                    const double* diff, double* out) {
    const int64_t R = p.rows, Cn = s.cols;
    for (int64_t r = 0; r < R; ++r) {
        double run = 0.0;
        for (int64_t c = 0; c < Cn; ++c) {
            run += diff ? diff[r * (Cn + 1) + c] : 0.0;
            out[c * R + r] = direct[r * Cn + c] + run;
        }
    }
    normalise_field(s, p, out);
}

void normalise_field(const HostScene& s, const HostProblem& p, double* sol) {
    const int64_t R = p.rows, Cn = s.cols;
    switch (p.tally) {
        case MCG_TEMP:
            for (int64_t i = 0; i < R * Cn; ++i) sol[i] = sol[i] / s.energy_sum;
            break;
        case MCG_MULTI:
            for (int64_t c = 0; c < Cn; ++c) sol[c * R] = sol[c * R] / s.energy_sum;
            break;
        case MCG_CUMTEMP:
            for (int64_t c = 0; c < Cn; ++c) {
                double* col = &sol[c * R];
                for (int64_t i = 0; i < p.size; ++i) col[i + 1] += col[i];
                for (int64_t r = 0; r < R; ++r) col[r] = col[r] / s.energy_sum;
            }
            break;
        case MCG_CUMFLUX:
            for (int64_t c = 0; c < Cn; ++c) {
                double* col = &sol[c * R];
                for (int64_t i = 0; i < p.size; ++i)
                    for (int k = 0; k < 3; ++k) col[3 * (i + 1) + k] += col[3 * i + k];
            }
            break;
        default:
            break;
    }
    for (int64_t c = 0; c < Cn; ++c)
        // a cell with no material (wholly inside a pore) reports zero
        for (int64_t r = 0; r < R; ++r)
            sol[c * R + r] = s.cell_vol[c] > 0.0 ? p.power * (sol[c * R + r] / s.cell_vol[c]) : 0.0;
}

}  // namespace mcg
