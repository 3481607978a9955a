/*
 * mcg_oracle.cpp -- FP64 CPU oracle of the phonon trajectory engine.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product links, loads or calls
 * this file: it is the checker that tests, the smoke check and the CPU
 * baseline leg of the benchmark compare the GPU engine against.
 *
 * It restates, in plain C++17 without Eigen or Boost, the reference's field
 * solve (nickdou/montecarlocpp, directory montecarlo/):
 *   Material::Material                 material.cpp:82-162
 *   Material::drawScatNext / scatter   material.cpp:215-231
 *   drawIso / drawAniso                random.cpp:16-44
 *   Phonon ctor / dir / move           phonon.cpp:33-36, 88-105
 *   Boundary planes and scatters       boundary.cpp:28-110, 147-187, 283-369,
 *                                      378-431, 456-460, 517-550
 *   Subdomain ctor / isInside / coord  subdomain.cpp:41-71, 108-116, 148-159
 *   Subdomain::advect                  subdomain.cpp:161-192
 *   emitter registry                   subdomain.cpp:199-212, domain.cpp:88-102
 *   EmitSubdomain, cell draws/volumes  subdomain.cpp:232-441
 *   Polygon faces                      boundary.cpp:190-251
 *   Field::init / accumulate           field.cpp:25-45, 92-220
 *   FieldProblem ctor / solve          problem.cpp:315-342, 370-445
 *   tally kinds                        problem.cpp:468-648
 *   solveField (OpenMP decomposition)  main.cpp:146-179
 *
 * Random numbers are NOT the reference's (Boost mt19937 seeded from
 * random_device, main.cpp:42-46).  Every bundle owns a Philox4x32-10
 * counter stream (Salmon, Moraes, Dror, Shaw, "Parallel Random Numbers: As
 * Easy as 1, 2, 3", SC'11), with a fixed assignment of words to draws per
 * event, so that a GPU engine following the same assignment reproduces every
 * trajectory.  Discrete draws use a Walker alias table (Walker 1977, built
 * with Vose's 1991 worklist method) over the (w, p) cells with positive
 * energy pdf; this has the distribution of the reference's two-level draw
 * (material.cpp:51-75): P(w, p) = pdf(w, p) / sum(pdf).
 *
 * Each part is pinned by an independent check in tests/: closed forms,
 * brute force on small cases, known-answer vectors, distribution tests and
 * invariants (the cell-volume formulas of the triangular prism and the
 * tetrahedron, subdomain.cpp:283-349, against brute-force cell shares in
 * test_geometry.py; the N-gon cells in test_ngon.py).
 */
#include "../include/mcg.h"
// The engine's portable FP64 log and sin/cos(pi r): the oracle uses them in
// place of libm so that CPU and GPU evaluate the same operations (see the
// header's comment); tests/test_math.py pins them against glibc.
#include "../montecarlocpp_amd/csrc/mcg_math.h"
// Rectangle-disk overlap for the cell volumes of pore-cut cells (pinned by
// tests/test_geometry.py against brute force).
#include "../montecarlocpp_amd/csrc/mcg_disk.h"

#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include <omp.h>

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

constexpr double HBAR = 1.054560652927e-034;     // constants.h:18
constexpr double KB = 1.380648e-023;             // constants.h:19

/* ------------------------------------------------------------ 3-vectors */
struct V3 {
    double x, y, z;
};
inline V3 v3(double a, double b, double c) { return V3{a, b, c}; }
inline V3 add(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
inline V3 sub(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline V3 scale(double s, V3 a) { return v3(s * a.x, s * a.y, s * a.z); }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) {
    return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
inline double norm(V3 a) { return std::sqrt(dot(a, a)); }
// Eigen's normalized(): divide by the norm when the squared norm is positive.
inline V3 normalized(V3 a) {
    double n2 = dot(a, a);
    if (n2 > 0.0) {
        double n = std::sqrt(n2);
        return v3(a.x / n, a.y / n, a.z / n);
    }
    return a;
}

/* 3x3 matrices, column-major: m[c*3 + r]. */
struct M3 {
    double m[9];
};
inline V3 col(const M3& a, int c) { return v3(a.m[3 * c], a.m[3 * c + 1], a.m[3 * c + 2]); }
inline V3 mul(const M3& a, V3 v) {
    return v3(a.m[0] * v.x + a.m[3] * v.y + a.m[6] * v.z,
              a.m[1] * v.x + a.m[4] * v.y + a.m[7] * v.z,
              a.m[2] * v.x + a.m[5] * v.y + a.m[8] * v.z);
}
inline M3 identity() {
    M3 r{};
    r.m[0] = r.m[4] = r.m[8] = 1.0;
    return r;
}
inline M3 transpose(const M3& a) {
    M3 r;
    for (int c = 0; c < 3; ++c)
        for (int rr = 0; rr < 3; ++rr) r.m[3 * c + rr] = a.m[3 * rr + c];
    return r;
}
inline double det(const M3& a) {
    // cofactor expansion along column 0
    double c00 = a.m[4] * a.m[8] - a.m[7] * a.m[5];
    double c10 = a.m[7] * a.m[2] - a.m[1] * a.m[8];
    double c20 = a.m[1] * a.m[5] - a.m[4] * a.m[2];
    return a.m[0] * c00 + a.m[3] * c10 + a.m[6] * c20;
}
// Inverse by the adjugate: inv = adj / det.
inline M3 inverse(const M3& a) {
    const double* m = a.m;
    double d = det(a);
    M3 r;
    // r(row, col) = cofactor(col, row) / det, with m(row, col) = m[col*3+row]
    r.m[0] = (m[4] * m[8] - m[7] * m[5]) / d;
    r.m[3] = (m[6] * m[5] - m[3] * m[8]) / d;
    r.m[6] = (m[3] * m[7] - m[6] * m[4]) / d;
    r.m[1] = (m[7] * m[2] - m[1] * m[8]) / d;
    r.m[4] = (m[0] * m[8] - m[6] * m[2]) / d;
    r.m[7] = (m[6] * m[1] - m[0] * m[7]) / d;
    r.m[2] = (m[1] * m[5] - m[4] * m[2]) / d;
    r.m[5] = (m[3] * m[2] - m[0] * m[5]) / d;
    r.m[8] = (m[0] * m[4] - m[3] * m[1]) / d;
    return r;
}
inline double frob(const M3& a) {
    double s = 0.0;
    for (double v : a.m) s += v * v;
    return std::sqrt(s);
}
// Eigen's isApprox for dense objects: ||a-b|| <= prec * min(||a||, ||b||).
inline bool mat_approx(const M3& a, const M3& b) {
    M3 d;
    for (int i = 0; i < 9; ++i) d.m[i] = a.m[i] - b.m[i];
    return frob(d) <= 1e-12 * std::min(frob(a), frob(b));
}
inline bool vec_approx(V3 a, V3 b) {
    return norm(sub(a, b)) <= 1e-12 * std::min(norm(a), norm(b));
}
// constants.h:21-29
inline bool scalar_approx(double a, double b) {
    double tol = 100 * DBL_EPSILON;
    double mx = std::max(std::fabs(a), std::fabs(b));
    double rhs = std::max(DBL_MIN, tol * mx);
    return std::fabs(a - b) <= rhs;
}

// Householder reflection I - 2 n n^T (boundary.cpp:28-31).
inline M3 refl_matrix(V3 n) {
    double nv[3] = {n.x, n.y, n.z};
    M3 r;
    for (int c = 0; c < 3; ++c)
        for (int rr = 0; rr < 3; ++rr)
            r.m[3 * c + rr] = (rr == c ? 1.0 : 0.0) - 2.0 * (nv[rr] * nv[c]);
    return r;
}

// Rotation taking the unit z axis onto n: the half-angle quaternion
// construction (boundary.cpp:33-37 uses Eigen's Quaternion::FromTwoVectors).
// For n antiparallel to z the rotation axis is taken to be x.
inline M3 rot_unit(V3 n);
inline M3 rot_matrix(V3 n_in) { return rot_unit(normalized(n_in)); }
inline M3 rot_unit(V3 n) {
    double c = n.z;
    double qw, qx, qy, qz;
    if (c < -1.0 + 1e-12) {
        c = std::max(c, -1.0);
        double w2 = (1.0 + c) * 0.5;
        qw = std::sqrt(w2);
        qx = std::sqrt(1.0 - w2);
        qy = 0.0;
        qz = 0.0;
    } else {
        // axis = z x n
        double ax = -n.y, ay = n.x, az = 0.0;
        double s = std::sqrt((1.0 + c) * 2.0);
        double invs = 1.0 / s;
        qx = ax * invs;
        qy = ay * invs;
        qz = az * invs;
        qw = s * 0.5;
    }
    double tx = 2.0 * qx, ty = 2.0 * qy, tz = 2.0 * qz;
    double twx = tx * qw, twy = ty * qw, twz = tz * qw;
    double txx = tx * qx, txy = ty * qx, txz = tz * qx;
    double tyy = ty * qy, tyz = tz * qy, tzz = tz * qz;
    M3 r;
    r.m[0] = 1.0 - (tyy + tzz);
    r.m[3] = txy - twz;
    r.m[6] = txz + twy;
    r.m[1] = txy + twz;
    r.m[4] = 1.0 - (txx + tzz);
    r.m[7] = tyz - twx;
    r.m[2] = txz - twy;
    r.m[5] = tyz + twx;
    r.m[8] = 1.0 - (txx + tyy);
    return r;
}

/* ------------------------------------------------------------ Philox */
// Philox4x32-10 (SC'11): 10 rounds of two 32x32->64 multiplies with the
// multipliers 0xD2511F53 / 0xCD9E8D57 and Weyl key increments 0x9E3779B9 /
// 0xBB67AE85.
inline void philox(const uint32_t ctr[4], uint32_t k0, uint32_t k1, uint32_t out[4]) {
    uint32_t x0 = ctr[0], x1 = ctr[1], x2 = ctr[2], x3 = ctr[3];
    for (int r = 0; r < 10; ++r) {
        if (r > 0) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        uint64_t p0 = uint64_t(0xD2511F53u) * x0;
        uint64_t p1 = uint64_t(0xCD9E8D57u) * x2;
        uint32_t y0 = uint32_t(p1 >> 32) ^ x1 ^ k0;
        uint32_t y1 = uint32_t(p1);
        uint32_t y2 = uint32_t(p0 >> 32) ^ x3 ^ k1;
        uint32_t y3 = uint32_t(p0);
        x0 = y0;
        x1 = y1;
        x2 = y2;
        x3 = y3;
    }
    out[0] = x0;
    out[1] = x1;
    out[2] = x2;
    out[3] = x3;
}

// Word slots of one event (emission = event 0, the scatter that follows
// loop iteration i = event i + 1).  A double uniform in slot w takes its high
// word from w and its low word from w + 4 (the same word of the next block);
// the integer draws take one word; free-path redraws take adjacent pairs.
enum : int {
    W_DIR1 = 0,     // direction: cos(theta) or sin^2(theta) draw
    W_DIR2 = 1,     // direction: azimuth draw
    W_PATH = 2,     // free path
    W_MODE = 3,     // (w, p) alias draw
    W_SPEC = 7,     // partial specularity test
    W_POS1 = 8,     // emission position draws
    W_POS2 = 9,
    W_POS3 = 10,
    W_PIECE = 11,   // emission: triangle of a polygon's fan, piece of an N-gon cell
    W_REDRAW = 16,  // free path redraw k: words 16 + 2k (high), 17 + 2k (low)
    B_POSRE = 64    // emission position redraw k: blocks 64 + 2k (high) and 65 + 2k (low)
};
// Cap on emission position redraws (a pore covers at most pi/4 of its cell's
// cross-section, so the cap is reached with probability below 1e-400).
constexpr int kMaxPosRedraw = 4096;

struct EventRng {
    uint32_t k0, k1, s, nlo, nhi;
    int cached[2] = {-1, -1};
    int next = 0;
    uint32_t buf[2][4];
    EventRng(uint64_t seed, uint32_t event, uint64_t bundle)
        : k0(uint32_t(seed)), k1(uint32_t(seed >> 32)), s(event),
          nlo(uint32_t(bundle)), nhi(uint32_t(bundle >> 32)) {}
    uint32_t word(int w) {
        const int j = w >> 2;
        for (int i = 0; i < 2; ++i)
            if (cached[i] == j) return buf[i][w & 3];
        const int i = next;   // two cached blocks, replaced in turn
        next ^= 1;
        uint32_t ctr[4] = {uint32_t(j), s, nlo, nhi};
        philox(ctr, k0, k1, buf[i]);
        cached[i] = j;
        return buf[i][w & 3];
    }
    // 53-bit uniform on [0, 1): 27 bits of word hi, 26 of word lo
    double u53_of(int hi, int lo) {
        uint32_t a = word(hi) >> 5;
        uint32_t b = word(lo) >> 6;
        return (double(a) * 67108864.0 + double(b)) * (1.0 / 9007199254740992.0);
    }
    double u53(int w) { return u53_of(w, w + 4); }
    // Boost uniform_real_distribution(-1, 1): u * (max - min) + min
    double upm1(int w) { return u53(w) * 2.0 + -1.0; }
};

/* --------------------------------------------------------- alias tables */
struct Alias {
    std::vector<uint32_t> thr, alias;
};

int build_alias(const double* pdf, int64_t n, Alias& out) {
    if (n <= 0) return fail(MCG_EINVAL, "empty distribution");
    double total = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        if (!(pdf[i] >= 0.0) || !std::isfinite(pdf[i]))
            return fail(MCG_EINVAL, "distribution has a negative or non-finite weight");
        total += pdf[i];
    }
    if (!(total > 0.0)) return fail(MCG_EINVAL, "distribution has zero total weight");
    std::vector<double> q(n);
    std::vector<int64_t> small, large;
    for (int64_t i = 0; i < n; ++i) {
        q[i] = pdf[i] * double(n) / total;
        if (q[i] < 1.0)
            small.push_back(i);
        else
            large.push_back(i);
    }
    std::vector<double> prob(n, 1.0);
    out.alias.assign(n, 0);
    for (int64_t i = 0; i < n; ++i) out.alias[i] = uint32_t(i);
    while (!small.empty() && !large.empty()) {
        int64_t s = small.back();
        small.pop_back();
        int64_t l = large.back();
        large.pop_back();
        prob[s] = q[s];
        out.alias[s] = uint32_t(l);
        q[l] = (q[l] + q[s]) - 1.0;
        if (q[l] < 1.0)
            small.push_back(l);
        else
            large.push_back(l);
    }
    out.thr.resize(n);
    for (int64_t i = 0; i < n; ++i) {
        double p = prob[i];
        if (p >= 1.0)
            out.thr[i] = 0xFFFFFFFFu;
        else
            out.thr[i] = uint32_t(p * 4294967296.0);
    }
    return MCG_OK;
}

inline uint32_t alias_draw(const Alias& t, uint32_t x) {
    uint64_t prod = uint64_t(x) * uint64_t(t.thr.size());
    uint32_t idx = uint32_t(prod >> 32);
    uint32_t frac = uint32_t(prod);
    return frac < t.thr[idx] ? idx : t.alias[idx];
}

/* ----------------------------------------------------------- material */
struct Material {
    int64_t nw = 0, np = 0;
    double T = 0;
    std::vector<double> omega, tau, vel, energy, flux, scat;
    double energySum = 0, fluxSum = 0, scatSum = 0, k = 0;
    std::vector<int64_t> modeCell;       // cell index w*np + p of each mode
    std::vector<double> modeVel, modeMfp;
    Alias fluxAlias, scatAlias;
};

int build_material(const mcg_material_spec* s, Material& m) {
    if (!s || !s->disp || !s->relax) return fail(MCG_EINVAL, "material spec is NULL");
    if (s->nw <= 0 || s->np <= 0) return fail(MCG_EINVAL, "Invalid dispersion file");
    if (!(s->temp > 0.0)) return fail(MCG_EINVAL, "temperature must be positive");
    const int64_t nw = s->nw, np = s->np, cols = 2 + 2 * np;
    m.nw = nw;
    m.np = np;
    m.T = s->temp;
    const double T = s->temp;
    m.omega.resize(nw);
    std::vector<double> domega(nw), dos(nw * np);
    m.vel.resize(nw * np);
    for (int64_t w = 0; w < nw; ++w) {
        const double* row = s->disp + w * cols;
        m.omega[w] = row[0];
        domega[w] = row[1];
        for (int64_t p = 0; p < np; ++p) {
            m.vel[w * np + p] = row[2 + 2 * p];
            dos[w * np + p] = row[3 + 2 * p];
        }
    }
    // tau (material.cpp:116-134)
    m.tau.assign(nw * np, 0.0);
    for (int64_t p = 0; p < np; ++p) {
        std::vector<double> tauinv(nw, 0.0);
        for (int j = 0; j < 2; ++j) {
            const double* c = s->relax + p * 8 + 4 * j;
            if (c[0] < 0.0) return fail(MCG_EINVAL, "Scattering times cannot be negative");
            if (c[0] <= DBL_MIN) continue;
            double tpow = std::pow(T, c[2]);
            double tex = std::exp(-c[3] / T);
            for (int64_t w = 0; w < nw; ++w)
                tauinv[w] = tauinv[w] + ((c[0] * std::pow(m.omega[w], c[1])) * tpow) * tex;
        }
        for (int64_t w = 0; w < nw; ++w) m.tau[w * np + p] = 1.0 / tauinv[w];
    }
    for (double t : m.tau)
        if (!std::isfinite(t))
            return fail(MCG_EINVAL, "Scattering time model produced infinite values");
    // dE/dT (material.cpp:136-145)
    std::vector<double> dedT(nw);
    const double xs = HBAR / (KB * T);
    for (int64_t w = 0; w < nw; ++w) {
        double x = xs * m.omega[w];
        double r;
        if (std::fabs(x) < DBL_EPSILON) {
            r = 1.0 - x * x / 12.0;
        } else {
            double q = x / (2.0 * std::sinh(x / 2.0));
            r = q * q;
        }
        dedT[w] = KB * r;
    }
    // pdfs and sums (material.cpp:147-161); sums run over w, then p
    m.energy.resize(nw * np);
    m.flux.resize(nw * np);
    m.scat.resize(nw * np);
    double ks = 0.0;
    for (int64_t w = 0; w < nw; ++w) {
        for (int64_t p = 0; p < np; ++p) {
            int64_t i = w * np + p;
            double e = (dedT[w] * dos[i]) * domega[w];
            if (e < 0.0) return fail(MCG_EINVAL, "negative density of states");
            m.energy[i] = e;
            m.flux[i] = m.vel[i] * e;
            m.scat[i] = e / m.tau[i];
            m.energySum += e;
            m.fluxSum += m.flux[i];
            m.scatSum += m.scat[i];
            ks += (m.tau[i] * (m.vel[i] * m.vel[i])) * e;
        }
    }
    m.k = ks / 3.0;
    // draw support: cells with positive energy pdf, w-major
    std::vector<double> fp, sp;
    for (int64_t i = 0; i < nw * np; ++i) {
        if (m.energy[i] > 0.0) {
            if ((m.flux[i] > 0.0 || m.scat[i] > 0.0) && !(m.vel[i] > DBL_MIN))
                return fail(MCG_EINVAL, "Velocity must be nonzero");
            m.modeCell.push_back(i);
            m.modeVel.push_back(m.vel[i]);
            m.modeMfp.push_back(m.vel[i] * m.tau[i]);
            fp.push_back(m.flux[i]);
            sp.push_back(m.scat[i]);
        }
    }
    if (m.modeCell.empty()) return fail(MCG_EINVAL, "material has no populated modes");
    int rc = build_alias(fp.data(), int64_t(fp.size()), m.fluxAlias);
    if (rc) return rc;
    rc = build_alias(sp.data(), int64_t(sp.size()), m.scatAlias);
    if (rc) return rc;
    return MCG_OK;
}

/* ----------------------------------------------------------- geometry */
enum ShapeKind { SHAPE_PAR = 0, SHAPE_TRI = 1, SHAPE_POLY = 2 };

// Discrete draw over a polygon's fan triangles or an N-gon cell's volume
// pieces: weights as the reference's DiscreteDist receives them, drawn by
// inverse CDF (the first k with u < cum[k]).
struct Pieces {
    int n = 0;
    V3 a[MCG_MAX_SIDES - 2], b[MCG_MAX_SIDES - 2];
    double cum[MCG_MAX_SIDES - 2];
    int pick(double u) const {
        int k = 0;
        while (k + 1 < n && !(u < cum[k])) ++k;
        return k;
    }
};

// Normalised cumulative weights; a weight below zero only by rounding is
// zero, and cum is exactly 1 from the last positive weight on.
int make_pieces(Pieces& p, const V3* a, const V3* b, const double* w, int n) {
    double mag = 0.0, total = 0.0, run = 0.0;
    int last = -1;
    for (int k = 0; k < n; ++k) mag += std::fabs(w[k]);
    for (int k = 0; k < n; ++k) {
        if (w[k] < -1e-12 * mag) return fail(MCG_EINVAL, "Volume too small, check vector order");
        if (w[k] > 0.0) {
            total += w[k];
            last = k;
        }
    }
    if (last < 0) return fail(MCG_EINVAL, "Volume too small, check vector order");
    p.n = n;
    for (int k = 0; k < n; ++k) {
        run += std::max(w[k], 0.0);
        p.cum[k] = k >= last ? 1.0 : run / total;
        p.a[k] = a[k];
        p.b[k] = b[k];
    }
    return MCG_OK;
}

struct Plane {
    V3 n;           // inward unit normal
    double c;       // offset: signed distance = n.p + c
    int kind;
    int sdom;
    // emitting faces (Isot, Peri)
    V3 o;           // face origin
    int shape;
    V3 si, sj;      // shape vectors
    Pieces fan;     // SHAPE_POLY: the fan's triangles (boundary.cpp:197-216, 244-251)
    double area;
    double T;       // after pairing
    bool emitter;   // registered at cell construction (subdomain.cpp:205-212)
    M3 rot;         // rotMatrix(normal) for Diff, Partial and emitters
    M3 refl;        // Spec, Partial
    uint32_t specThr;
    // Peri
    int pair = -1;
    M3 periRot;
    V3 periT;
    // Inter
    std::vector<int> pairs;
};

struct Sdom {
    int pore = -1;
    int cell;
    V3 o;
    M3 mat, inv;
    int64_t div[3], shape[3], max[3];
    int accum;
    double eps;
    double vol;
    V3 gradT;
    M3 rotG;
    Pieces vpieces; // prism / pyramid volume pieces (subdomain.cpp:385-441)
    V3 axis;        // their column 0: the prism's axis, the pyramid's apex
    int pbeg, pend;
    int64_t colOff;
    int64_t stride[3];
    int64_t ncell;
};

struct Emitter {
    bool volume;
    int index;     // sdom (volume) or plane
    double weight;
};

// Extension: a cylindrical pore (axis along z) cut out of a box cell.
struct Pore {
    int sdom, kind;
    double cx, cy, r, r2, eps;
    uint32_t thr;
};

struct Domain {
    std::vector<Pore> pores;
    std::vector<Plane> planes;
    std::vector<Sdom> sdoms;
    std::vector<Emitter> emitters;
    int64_t cols = 0;
    std::vector<double> cellVol;
};

struct FaceGeo {
    V3 o, i, j;
    int shape;
    std::vector<V3> fan;   // SHAPE_POLY: vertices after the first, relative to o
};

// Sides N of an N-gon cell's base from its face count (0 when not in 4..9).
int ngon_sides(const mcg_subdomain_spec& s) {
    const int n = s.cell == MCG_PRISM ? s.nface - 2 : s.cell == MCG_PYRAMID ? s.nface - 1 : 0;
    return n >= 4 && n <= MCG_MAX_SIDES ? n : 0;
}

V3 mat_col(const mcg_subdomain_spec& s, int k) {
    return v3(s.mat[3 * k], s.mat[3 * k + 1], s.mat[3 * k + 2]);
}

// PrismImpl / PyramidImpl::volume (subdomain.cpp:385-394, 417-426), as
// written: term i = (m_i x m_(i+1)) . m_0 / div for i = 0..N-3.
void ngon_volume_terms(const mcg_subdomain_spec& s, int n, double div, double* w) {
    const V3 a = mat_col(s, 0);
    for (int i = 0; i < n - 2; ++i) w[i] = dot(cross(mat_col(s, i), mat_col(s, i + 1)), a) / div;
}

int cell_faces(const mcg_subdomain_spec& s, const M3& mat, V3 o, std::vector<FaceGeo>& f) {
    V3 m0 = col(mat, 0), m1 = col(mat, 1), m2 = col(mat, 2);
    f.clear();
    switch (s.cell) {
        case MCG_PARALLELEPIPED:  // subdomain.h:149-154
            f.push_back({o, m1, m2, SHAPE_PAR, {}});
            f.push_back({o, m2, m0, SHAPE_PAR, {}});
            f.push_back({o, m0, m1, SHAPE_PAR, {}});
            f.push_back({add(o, m0), m2, m1, SHAPE_PAR, {}});
            f.push_back({add(o, m1), m0, m2, SHAPE_PAR, {}});
            f.push_back({add(o, m2), m1, m0, SHAPE_PAR, {}});
            return MCG_OK;
        case MCG_TRIPRISM:  // subdomain.h:228-233
            f.push_back({o, m1, m2, SHAPE_PAR, {}});
            f.push_back({o, m2, m0, SHAPE_PAR, {}});
            f.push_back({o, m0, m1, SHAPE_TRI, {}});
            f.push_back({add(o, m0), m2, sub(m1, m0), SHAPE_PAR, {}});
            f.push_back({add(o, m2), m1, m0, SHAPE_TRI, {}});
            return MCG_OK;
        case MCG_TETRAHEDRON:  // subdomain.h:305-309
            f.push_back({o, m1, m2, SHAPE_TRI, {}});
            f.push_back({o, m2, m0, SHAPE_TRI, {}});
            f.push_back({o, m0, m1, SHAPE_TRI, {}});
            f.push_back({add(o, m0), sub(m2, m0), sub(m1, m0), SHAPE_TRI, {}});
            return MCG_OK;
        case MCG_PRISM:     // subdomain.h:398-409 (Bot, Top), 462-479 (sides)
        case MCG_PYRAMID: { // subdomain.h:537-546 (Bot), 598-617 (sides)
            const int n = ngon_sides(s);
            if (!n) return fail(MCG_EINVAL, "a prism or pyramid base needs 4 to 9 sides");
            std::vector<V3> bot, top;
            for (int k = 1; k < n; ++k) bot.push_back(mat_col(s, k));
            for (int k = n - 1; k >= 1; --k) top.push_back(mat_col(s, k));
            const V3 ax = mat_col(s, 0);
            f.push_back({o, bot[0], bot[1], SHAPE_POLY, bot});
            if (s.cell == MCG_PRISM) f.push_back({add(o, ax), top[0], top[1], SHAPE_POLY, top});
            for (int ind = 0; ind < n; ++ind) {
                V3 p = v3(0.0, 0.0, 0.0);
                if (ind != 0) p = add(p, mat_col(s, ind));
                V3 j = scale(-1.0, p);
                if (ind != n - 1) j = add(j, mat_col(s, ind + 1));
                if (s.cell == MCG_PRISM)
                    f.push_back({add(o, p), ax, j, SHAPE_PAR, {}});
                else
                    f.push_back({add(o, p), sub(ax, p), j, SHAPE_TRI, {}});
            }
            return MCG_OK;
        }
    }
    return fail(MCG_EINVAL, "unknown cell kind");
}

double cell_volume(const Sdom& s, const int64_t idx[3]) {
    switch (s.cell) {
        case MCG_PARALLELEPIPED:  // subdomain.cpp:269-273
            return s.vol / double(s.shape[0] * s.shape[1] * s.shape[2]);
        case MCG_PRISM:           // subdomain.cpp:396-399
        case MCG_PYRAMID:         // subdomain.cpp:428-431
            return s.vol;
        case MCG_TRIPRISM: {  // subdomain.cpp:283-307
            double s0 = double(s.shape[0]), s1 = double(s.shape[1]);
            double f0 = 1.0 - (double(idx[0]) / s0 + double(idx[1]) / s1);
            if (f0 <= 0.0) return 0.0;
            double f1 = f0 - (1.0 / s0 + 1.0 / s1);
            if (f1 >= 0.0) return s.vol / double(s.shape[0] * s.shape[1] * s.shape[2]);
            double frac = f0 * f0;
            const int pts[2][2] = {{1, 0}, {0, 1}};
            for (int i = 0; i < 2; ++i) {
                double f = f0 - (pts[i][0] / s0 + pts[i][1] / s1);
                if (f > 0.0) frac += -1.0 * (f * f);
            }
            return s.vol * frac / (2.0 * double(s.shape[2]));
        }
        case MCG_TETRAHEDRON: {  // subdomain.cpp:322-349
            double s0 = double(s.shape[0]), s1 = double(s.shape[1]), s2 = double(s.shape[2]);
            double f0 = 1.0 - (double(idx[0]) / s0 + double(idx[1]) / s1 + double(idx[2]) / s2);
            if (f0 <= 0.0) return 0.0;
            double f1 = f0 - (1.0 / s0 + 1.0 / s1 + 1.0 / s2);
            if (f1 >= 0.0) return s.vol / double(s.shape[0] * s.shape[1] * s.shape[2]);
            double frac = f0 * f0 * f0;
            const int pts[6][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {0, 1, 1}, {1, 0, 1}, {1, 1, 0}};
            for (int i = 0; i < 6; ++i) {
                int sum = pts[i][0] + pts[i][1] + pts[i][2];
                double sign = (sum % 2 == 0) ? 1.0 : -1.0;
                double f = f0 - (pts[i][0] / s0 + pts[i][1] / s1 + pts[i][2] / s2);
                if (f > 0.0) frac += sign * (f * f * f);
            }
            return s.vol * frac / 6.0;
        }
    }
    return 0.0;
}

int build_domain(const mcg_domain_spec* d, Domain& D) {
    if (!d || d->nsdom <= 0 || !d->sdom) return fail(MCG_EINVAL, "Domain setup not complete");
    D.pores.clear();
    D.planes.clear();
    D.sdoms.clear();
    D.emitters.clear();
    std::vector<std::vector<int>> faceEmit(d->nsdom);
    for (int si = 0; si < d->nsdom; ++si) {
        const mcg_subdomain_spec& s = d->sdom[si];
        Sdom S{};
        S.cell = s.cell;
        S.o = v3(s.origin[0], s.origin[1], s.origin[2]);
        std::memcpy(S.mat.m, s.mat, sizeof(S.mat.m));
        if (s.cell < MCG_PARALLELEPIPED || s.cell > MCG_PYRAMID) return fail(MCG_EINVAL, "unknown cell kind");
        const bool ngon = s.cell == MCG_PRISM || s.cell == MCG_PYRAMID;
        const int nside = ngon_sides(s);
        if (ngon && !nside) return fail(MCG_EINVAL, "a prism or pyramid base needs 4 to 9 sides");
        int nface_expect = s.cell == MCG_PARALLELEPIPED ? 6 : s.cell == MCG_TRIPRISM ? 5
                         : s.cell == MCG_TETRAHEDRON ? 4 : s.nface;
        if (s.nface != nface_expect) return fail(MCG_EINVAL, "Incorrect number of temperatures");
        if (ngon) {
            // the cell's coordinate matrix (m_1, m_(N-1), m_0) (subdomain.cpp:379-383, 411-415)
            const V3 c[3] = {mat_col(s, 1), mat_col(s, nside - 1), mat_col(s, 0)};
            for (int k = 0; k < 3; ++k) {
                S.mat.m[3 * k] = c[k].x;
                S.mat.m[3 * k + 1] = c[k].y;
                S.mat.m[3 * k + 2] = c[k].z;
            }
            double w[MCG_MAX_SIDES];
            ngon_volume_terms(s, nside, s.cell == MCG_PRISM ? 2.0 : 6.0, w);
            S.vol = 0.0;
            for (int k = 0; k < nside - 2; ++k) S.vol += w[k];
        } else {
            double dt = det(S.mat);
            S.vol = s.cell == MCG_PARALLELEPIPED ? dt : s.cell == MCG_TRIPRISM ? dt / 2.0 : dt / 6.0;
        }
        if (!(S.vol >= DBL_MIN)) return fail(MCG_EINVAL, "Volume too small, check vector order");
        S.inv = inverse(S.mat);
        // subdomain.cpp:41-71; an N-gon cell's div is (-1, -1, -1) or
        // (0, 0, 0) by the sign of its one div (subdomain.h:388-391)
        bool anyNeg = false;
        int npos = 0;
        int64_t dv[3];
        for (int k = 0; k < 3; ++k) dv[k] = ngon ? (s.div[0] < 0 ? -1 : 0) : s.div[k];
        for (int k = 0; k < 3; ++k) {
            S.div[k] = dv[k];
            S.shape[k] = std::max<int64_t>(dv[k], 1);
            S.max[k] = S.shape[k] - 1;
            if (dv[k] < 0) anyNeg = true;
            if (dv[k] > 0) ++npos;
        }
        if (anyNeg) {
            for (int k = 0; k < 3; ++k) S.shape[k] = S.max[k] = 0;
            S.accum = -2;
        } else if (npos == 0) {
            S.accum = -1;
        } else if (npos == 1) {
            int dir = 0;
            for (int k = 1; k < 3; ++k)
                if (dv[k] > dv[dir]) dir = k;
            S.accum = dir;
        } else {
            S.accum = npos + 1;
        }
        if (ngon) {
            // drawPos: piece i has edges (m_(i+1), m_(i+2), m_0) and weight
            // volume term i (subdomain.cpp:401-409, 433-441)
            double w[MCG_MAX_SIDES];
            ngon_volume_terms(s, nside, s.cell == MCG_PRISM ? 2.0 : 6.0, w);
            V3 pa[MCG_MAX_SIDES], pb[MCG_MAX_SIDES];
            for (int k = 0; k < nside - 2; ++k) {
                pa[k] = mat_col(s, k + 1);
                pb[k] = mat_col(s, k + 2);
            }
            S.axis = mat_col(s, 0);
            if (s.grad_t[0] != 0.0 || s.grad_t[1] != 0.0 || s.grad_t[2] != 0.0) {
                int rc = make_pieces(S.vpieces, pa, pb, w, nside - 2);
                if (rc) return rc;
            }
        }
        double mx = 0.0;
        for (int k = 0; k < 3; ++k) mx = std::max(mx, norm(col(S.mat, k)));
        S.eps = 100.0 * DBL_EPSILON * mx;
        S.gradT = v3(s.grad_t[0], s.grad_t[1], s.grad_t[2]);
        S.rotG = rot_matrix(normalized(S.gradT));
        std::vector<FaceGeo> fg;
        int rc = cell_faces(s, S.mat, S.o, fg);
        if (rc) return rc;
        S.pbeg = int(D.planes.size());
        for (int f = 0; f < s.nface; ++f) {
            const mcg_face_spec& fs = s.face[f];
            Plane P{};
            P.kind = fs.kind;
            P.sdom = si;
            P.o = fg[f].o;
            P.shape = fg[f].shape;
            P.si = fg[f].i;
            P.sj = fg[f].j;
            V3 cr = cross(P.si, P.sj);
            P.area = P.shape == SHAPE_PAR ? norm(cr) : norm(cr) / 2.0;
            P.n = normalized(cr);
            if (P.shape == SHAPE_POLY) {
                // Polygon<N>: fan triangles of positive area, all facing the
                // polygon's normal; area is their sum (boundary.cpp:197-216, 238-242)
                const std::vector<V3>& v = fg[f].fan;
                double w[MCG_MAX_SIDES];
                P.area = 0.0;
                for (size_t k = 0; k + 1 < v.size(); ++k) {
                    const V3 t = cross(v[k], v[k + 1]);
                    w[k] = norm(t) / 2.0;
                    if (!(w[k] > DBL_MIN)) return fail(MCG_EINVAL, "Area too small");
                    if (!vec_approx(normalized(t), P.n)) return fail(MCG_EINVAL, "Normals are inconsistent");
                    P.area += w[k];
                }
                int rc = make_pieces(P.fan, v.data(), v.data() + 1, w, int(v.size()) - 1);
                if (rc) return rc;
            }
            if (P.area == 0.0) return fail(MCG_EINVAL, "Shape not initialized");
            P.c = -dot(P.n, P.o);
            P.T = fs.temp;
            if (fs.kind < MCG_SPEC || fs.kind > MCG_PARTIAL)
                return fail(MCG_EINVAL, "unknown boundary kind");
            bool emitting = fs.kind == MCG_ISOT || fs.kind == MCG_PERI;
            if (!emitting && fs.temp != 0.0) return fail(MCG_EINVAL, "Cannot specify temperature");
            P.rot = rot_matrix(P.n);
            P.refl = refl_matrix(P.n);
            if (fs.kind == MCG_PARTIAL) {
                double p = fs.specularity;
                if (!(p >= 0.0 && p <= 1.0)) return fail(MCG_EINVAL, "specularity outside [0, 1]");
                P.specThr = p >= 1.0 ? 0xFFFFFFFFu : uint32_t(p * 4294967296.0);
                // the end points are SpecBoundary and DiffBoundary themselves
                // (boundary.cpp:283-287, 308-312): no draw is taken there
                if (p >= 1.0) P.kind = MCG_SPEC;
                if (p <= 0.0) P.kind = MCG_DIFF;
            }
            P.emitter = emitting && (P.area * std::fabs(P.T) != 0.0);
            if (P.emitter) faceEmit[si].push_back(int(D.planes.size()));
            D.planes.push_back(P);
        }
        S.pend = int(D.planes.size());
        D.sdoms.push_back(S);
    }
    // pairs (boundary.cpp:361-369, 524-550)
    for (int k = 0; k < d->npair; ++k) {
        const mcg_pair_spec& ps = d->pair[k];
        if (ps.sdom_a < 0 || ps.sdom_a >= d->nsdom || ps.sdom_b < 0 || ps.sdom_b >= d->nsdom)
            return fail(MCG_EINVAL, "pair names an unknown subdomain");
        const Sdom& A = D.sdoms[ps.sdom_a];
        const Sdom& B = D.sdoms[ps.sdom_b];
        if (ps.face_a < 0 || ps.face_a >= A.pend - A.pbeg || ps.face_b < 0 ||
            ps.face_b >= B.pend - B.pbeg)
            return fail(MCG_EINVAL, "pair names an unknown face");
        int ia = A.pbeg + ps.face_a, ib = B.pbeg + ps.face_b;
        Plane& a = D.planes[ia];
        Plane& b = D.planes[ib];
        if (a.kind == MCG_PERI && b.kind == MCG_PERI) {
            M3 R;
            bool zero = true;
            for (double v : ps.rot) zero = zero && v == 0.0;
            if (zero)
                R = identity();
            else
                std::memcpy(R.m, ps.rot, sizeof(R.m));
            V3 t = v3(ps.transl[0], ps.transl[1], ps.transl[2]);
            if (a.pair >= 0 || b.pair >= 0) return fail(MCG_EINVAL, "Boundary already paired");
            M3 RRt;
            M3 Rt = transpose(R);
            for (int c = 0; c < 3; ++c) {
                V3 cc = mul(R, col(Rt, c));
                RRt.m[3 * c] = cc.x;
                RRt.m[3 * c + 1] = cc.y;
                RRt.m[3 * c + 2] = cc.z;
            }
            if (!mat_approx(identity(), RRt))
                return fail(MCG_EINVAL, "Rotation matrix must be orthogonal");
            if (!vec_approx(mul(R, a.n), scale(-1.0, b.n)))
                return fail(MCG_EINVAL, "Rotation matrix incorrect");
            V3 tr = add(scale(-a.c, mul(R, a.n)), t);
            if (!scalar_approx(dot(b.n, tr), -b.c))
                return fail(MCG_EINVAL, "Translation vector incorrect");
            double T1 = a.T, T2 = b.T;
            a.T = T1 - T2;
            b.T = T2 - T1;
            a.periRot = R;
            b.periRot = Rt;
            a.periT = t;
            b.periT = scale(-1.0, mul(Rt, t));
            a.pair = ib;
            b.pair = ia;
        } else if (a.kind == MCG_INTER && b.kind == MCG_INTER) {
            if (!vec_approx(a.n, scale(-1.0, b.n)))
                return fail(MCG_EINVAL, "Boundary normals not antiparallel");
            if (!scalar_approx(a.c, -b.c)) return fail(MCG_EINVAL, "Boundaries planes not the same");
            a.pairs.push_back(ib);
            b.pairs.push_back(ia);
        } else {
            return fail(MCG_EINVAL, "pair joins faces that are not both Peri or both Inter");
        }
    }
    for (const Plane& P : D.planes) {
        if (P.kind == MCG_PERI && P.pair < 0) return fail(MCG_EINVAL, "Domain setup not complete");
        if (P.kind == MCG_INTER && P.pairs.empty())
            return fail(MCG_EINVAL, "Domain setup not complete");
    }
    // pores: axis-aligned box cells only, disk inside the cell's x-y extent
    for (int k = 0; k < d->npore; ++k) {
        const mcg_pore_spec& ps = d->pore[k];
        if (ps.sdom < 0 || ps.sdom >= d->nsdom) return fail(MCG_EINVAL, "pore names an unknown subdomain");
        Sdom& S = D.sdoms[ps.sdom];
        if (S.pore >= 0) return fail(MCG_EINVAL, "one pore per subdomain");
        const double* m = S.mat.m;
        if (S.cell != MCG_PARALLELEPIPED || m[1] != 0.0 || m[2] != 0.0 || m[3] != 0.0 || m[5] != 0.0 ||
            m[6] != 0.0 || m[7] != 0.0)
            return fail(MCG_EINVAL, "a pore needs an axis-aligned box cell");
        if (!(ps.radius > 0.0)) return fail(MCG_EINVAL, "pore radius must be positive");
        if (ps.center[0] - ps.radius < S.o.x || ps.center[0] + ps.radius > S.o.x + m[0] ||
            ps.center[1] - ps.radius < S.o.y || ps.center[1] + ps.radius > S.o.y + m[4])
            return fail(MCG_EINVAL, "pore must lie inside the cell");
        if (ps.kind != MCG_DIFF && ps.kind != MCG_SPEC && ps.kind != MCG_PARTIAL)
            return fail(MCG_EINVAL, "pore wall must be Diff, Spec or Partial");
        Pore P{ps.sdom, ps.kind, ps.center[0], ps.center[1], ps.radius, ps.radius * ps.radius, S.eps, 0u};
        if (ps.kind == MCG_PARTIAL) {
            if (!(ps.specularity >= 0.0 && ps.specularity <= 1.0))
                return fail(MCG_EINVAL, "specularity outside [0, 1]");
            P.thr = ps.specularity >= 1.0 ? 0xFFFFFFFFu : uint32_t(ps.specularity * 4294967296.0);
            if (ps.specularity >= 1.0) P.kind = MCG_SPEC;
            if (ps.specularity <= 0.0) P.kind = MCG_DIFF;
        }
        // the pore runs through the cell's faces across its axis: a face
        // there keeps the hole (Peri back onto this cell) or closes it; an
        // Inter face, or a Peri face onto another cell, would let bundles in
        // through the hole
        for (int b = S.pbeg; b < S.pend; ++b) {
            const Plane& F = D.planes[b];
            if (F.n.z == 0.0) continue;
            if (F.kind == MCG_INTER) return fail(MCG_EINVAL, "a pore cannot open onto an Inter face");
            if (F.kind == MCG_PERI && D.planes[F.pair].sdom != ps.sdom)
                return fail(MCG_EINVAL, "a pore's Peri faces must pair within its cell");
        }
        S.pore = int(D.pores.size());
        D.pores.push_back(P);
    }
    // emitter registry (domain.cpp:88-102, subdomain.cpp:205-212); a pore
    // removes its volume from the gradient source
    for (int si = 0; si < d->nsdom; ++si) {
        Sdom& S = D.sdoms[si];
        double vmat = S.vol;
        if (S.pore >= 0) {
            const Pore& P = D.pores[S.pore];
            vmat = S.vol - 3.141592653589793 * (P.r * P.r) * S.mat.m[8];
        }
        double wv = 2.0 * vmat * norm(S.gradT);
        if (wv != 0.0) D.emitters.push_back({true, si, wv});
        for (int pi : faceEmit[si]) {
            const Plane& P = D.planes[pi];
            double area = P.area;   // a face across the pore's axis has its disk cut out
            if (S.pore >= 0 && P.n.z != 0.0) area -= 3.141592653589793 * (D.pores[S.pore].r * D.pores[S.pore].r);
            D.emitters.push_back({false, pi, area * std::fabs(P.T)});
        }
    }
    // field columns (field.cpp:25-45) and cell volumes (field.cpp:53-77)
    D.cols = 0;
    D.cellVol.clear();
    for (Sdom& S : D.sdoms) {
        int64_t prod = S.shape[0] * S.shape[1] * S.shape[2];
        S.ncell = prod;
        S.colOff = D.cols;
        S.stride[0] = 1;
        S.stride[1] = S.shape[0];
        S.stride[2] = S.shape[0] * S.shape[1];
        if (prod == 0) continue;
        for (int64_t k = 0; k < S.shape[2]; ++k)
            for (int64_t j = 0; j < S.shape[1]; ++j)
                for (int64_t i = 0; i < S.shape[0]; ++i) {
                    int64_t idx[3] = {i, j, k};
                    double v = cell_volume(S, idx);
                    if (S.pore >= 0) {
                        const Pore& P = D.pores[S.pore];
                        const double lx = S.mat.m[0], ly = S.mat.m[4];
                        const double x0 = S.o.x + lx * (double(i) / double(S.shape[0]));
                        const double x1 = S.o.x + lx * (double(i + 1) / double(S.shape[0]));
                        const double y0 = S.o.y + ly * (double(j) / double(S.shape[1]));
                        const double y1 = S.o.y + ly * (double(j + 1) / double(S.shape[1]));
                        const double in = mcg_disk::rect_in_disk(x0, x1, y0, y1, P.cx, P.cy, P.r);
                        v = v * (1.0 - in / ((x1 - x0) * (y1 - y0)));
                    }
                    D.cellVol.push_back(v);
                }
        D.cols += prod;
    }
    return MCG_OK;
}

/* ------------------------------------------------------------ problem */
struct Problem {
    int tally;
    int64_t rows;
    int64_t nemit, maxscat, maxloop, size, step;
    double power, weightSum;
    std::vector<int64_t> emitPdf, emitCdf;
    uint64_t seed;
    int64_t first, count;
};

// trace_only: a trajectory from a given start (TrajProblem with pos_ set,
// problem.cpp:226-299) never touches the emitters, so a domain without a
// source is accepted and nemit, counts and power stay empty
int build_problem(const mcg_problem_spec* p, const Material& M, const Domain& D, Problem& P,
                  bool trace_only = false) {
    if (!p) return fail(MCG_EINVAL, "problem spec is NULL");
    if (p->tally < MCG_TEMP || p->tally > MCG_CUMFLUX) return fail(MCG_EINVAL, "Invalid problem");
    if (p->nemit <= 0) return fail(MCG_EINVAL, "nemit must be positive");
    if (p->maxscat <= 0) return fail(MCG_EINVAL, "maxscat must be positive");
    if (p->maxloop < 0) return fail(MCG_EINVAL, "maxloop must be non-negative");
    if (p->maxloop == 0 && p->maxscat > INT64_MAX / 100)
        return fail(MCG_ERANGE, "maxscat too large for the default maxloop (100 maxscat)");
    if (D.emitters.empty() && !trace_only) return fail(MCG_EINVAL, "No phonons to emit");
    P.tally = p->tally;
    P.size = p->size;
    switch (p->tally) {  // problem.cpp:468-471, 501-504, 534-537, 576-579, 624-627
        case MCG_TEMP: P.rows = 1; break;
        case MCG_FLUX: P.rows = 3; break;
        case MCG_MULTI: P.rows = 4; break;
        case MCG_CUMTEMP: P.rows = p->size + 1; break;
        case MCG_CUMFLUX: P.rows = 3 * (p->size + 1); break;
    }
    P.step = 0;
    if (p->tally == MCG_CUMTEMP || p->tally == MCG_CUMFLUX) {
        if (p->size <= 0) return fail(MCG_EINVAL, "size must be positive");
        // problem.cpp:562-563
        P.step = (p->maxscat - 1) / p->size;
        if ((p->maxscat - 1) % p->size != 0) P.step++;
        if (P.step <= 0) return fail(MCG_EINVAL, "maxscat must exceed 1 for a cumulative tally");
    }
    P.maxscat = p->maxscat;
    P.maxloop = p->maxloop != 0 ? p->maxloop : 100 * p->maxscat;
    P.seed = p->seed;
    if (trace_only) {
        P.weightSum = P.power = 0.0;
        P.emitPdf.clear();
        P.emitCdf.clear();
        P.nemit = p->nemit;
        P.first = p->first_bundle;
        P.count = 1;
        return MCG_OK;
    }
    // problem.cpp:319-341
    double wsum = 0.0;
    for (const Emitter& e : D.emitters) {
        if (!(e.weight > 0.0)) return fail(MCG_EINVAL, "No phonons to emit");
        wsum += e.weight;
    }
    P.weightSum = wsum;
    P.emitPdf.clear();
    for (const Emitter& e : D.emitters) {
        double frac = e.weight / wsum;
        double rounded = std::ceil(frac * double(p->nemit) - 0.5);
        P.emitPdf.push_back(std::max<int64_t>(1, int64_t(rounded)));
    }
    P.nemit = 0;
    P.emitCdf.clear();
    for (int64_t c : P.emitPdf) {
        P.nemit += c;
        P.emitCdf.push_back(P.nemit);
    }
    P.power = wsum / double(P.nemit) * M.fluxSum / 4.0;
    P.first = p->first_bundle;
    P.count = p->bundle_count == 0 ? P.nemit - P.first : p->bundle_count;
    if (P.first < 0 || P.count < 0 || P.first + P.count > P.nemit)
        return fail(MCG_ERANGE, "bundle range outside [0, nemit)");
    return MCG_OK;
}

/* ------------------------------------------------------------ phonon */
struct Phonon {
    bool alive;
    int sign;
    uint32_t mode;
    V3 pos, dir;
    double scatNext;
    int64_t nscat;
    // TrkPhonon (phonon.cpp:133-157): when set, every new position is
    // appended, as its pos(newPos) does
    std::vector<V3>* track;
};

// Phonon::pos(newPos) / TrkPhonon::pos(newPos) (phonon.cpp:154-157)
inline void set_pos(Phonon& ph, const V3& p) {
    ph.pos = p;
    if (ph.track) ph.track->push_back(p);
}

// random.cpp:16-27
V3 draw_iso(EventRng& g) {
    double cosTheta = g.upm1(W_DIR1);
    double sinTheta = std::sqrt(1.0 - cosTheta * cosTheta);
    double sphi, cphi;   // phi = pi * U[-1, 1)
    mcg_math::sincospi(g.upm1(W_DIR2), &sphi, &cphi);
    return v3(sinTheta * cphi, sinTheta * sphi, cosTheta);
}

// random.cpp:29-44
V3 draw_aniso(EventRng& g, bool bidir) {
    double r = g.upm1(W_DIR1);
    int sign = bidir ? (r < 0.0 ? -1 : 1) : 1;
    double sinSqTheta = std::fabs(r);
    double sinTheta = std::sqrt(sinSqTheta);
    double cosTheta = double(sign) * std::sqrt(1.0 - sinSqTheta);
    double sphi, cphi;   // phi = pi * U[-1, 1)
    mcg_math::sincospi(g.upm1(W_DIR2), &sphi, &cphi);
    return v3(sinTheta * cphi, sinTheta * sphi, cosTheta);
}

// material.cpp:215-224 with phonon.cpp:122-127
void draw_scat_next(const Material& M, Phonon& ph, EventRng& g) {
    double mfp = M.modeMfp[ph.mode];
    double distance = mfp * -mcg_math::log(1.0 - g.u53(W_PATH));
    for (int k = 0; distance < DBL_MIN; ++k)
        distance = mfp * -mcg_math::log(1.0 - g.u53_of(W_REDRAW + 2 * k, W_REDRAW + 2 * k + 1));
    ph.scatNext = distance;
}

// Emitter::emit (boundary.cpp:378-385) with the wall (boundary.cpp:418-431,
// 147-152, 182-187) and volume (subdomain.cpp:255-263, 275-377) draws.
Phonon emit(const Material& M, const Domain& D, const Emitter& e, EventRng& g, int& sdom) {
    Phonon ph{};
    ph.mode = alias_draw(M.fluxAlias, g.word(W_MODE));
    V3 pos, dir;
    bool sign;
    if (!e.volume) {
        const Plane& P = D.planes[e.index];
        double r1 = g.u53(W_POS1), r2 = g.u53(W_POS2);
        V3 si = P.si, sj = P.sj;
        if (P.shape == SHAPE_POLY) {   // Polygon<N>::drawPos (boundary.cpp:244-251)
            const int k = P.fan.pick(g.u53(W_PIECE));
            si = P.fan.a[k];
            sj = P.fan.b[k];
        }
        V3 local;
        if (P.shape == SHAPE_PAR || r1 + r2 < 1.0)
            local = add(scale(r1, si), scale(r2, sj));
        else
            local = add(scale(1.0 - r1, si), scale(1.0 - r2, sj));
        pos = add(P.o, local);
        if (D.sdoms[P.sdom].pore >= 0) {   // a wall across the pore: redraw inside its disk
            const Pore& Q = D.pores[D.sdoms[P.sdom].pore];
            for (int k = 0; k < kMaxPosRedraw; ++k) {
                const double dx = pos.x - Q.cx, dy = pos.y - Q.cy;
                if (!(dx * dx + dy * dy < Q.r2)) break;
                const int w = 4 * (B_POSRE + 2 * k);
                pos = add(P.o, add(scale(g.u53(w), P.si), scale(g.u53(w + 1), P.sj)));
            }
        }
        dir = mul(P.rot, draw_aniso(g, false));
        sign = P.T >= 0.0;
        sdom = P.sdom;
    } else {
        const Sdom& S = D.sdoms[e.index];
        double c[3] = {g.u53(W_POS1), g.u53(W_POS2), g.u53(W_POS3)};
        if (S.cell != MCG_PARALLELEPIPED) {
            if (c[0] + c[1] > 1.0) {
                c[0] = 1.0 - c[0];
                c[1] = 1.0 - c[1];
            }
        }
        if (S.cell == MCG_TETRAHEDRON || S.cell == MCG_PYRAMID) {
            if (c[1] + c[2] > 1.0) {
                double tmp = c[2];
                c[2] = 1.0 - c[0] - c[1];
                c[1] = 1.0 - tmp;
            } else if (c[0] + c[1] + c[2] > 1.0) {
                double tmp = c[2];
                c[2] = c[0] + c[1] + c[2] - 1.0;
                c[0] = 1.0 - c[1] - tmp;
            }
        }
        M3 local = S.mat;
        if (S.cell == MCG_PRISM || S.cell == MCG_PYRAMID) {
            // Prism / PyramidImpl::drawPos (subdomain.cpp:401-409, 433-441)
            const int k = S.vpieces.pick(g.u53(W_PIECE));
            const V3 cols[3] = {S.vpieces.a[k], S.vpieces.b[k], S.axis};
            for (int q = 0; q < 3; ++q) {
                local.m[3 * q] = cols[q].x;
                local.m[3 * q + 1] = cols[q].y;
                local.m[3 * q + 2] = cols[q].z;
            }
        }
        pos = add(S.o, mul(local, v3(c[0], c[1], c[2])));
        if (S.pore >= 0) {   // uniform in the box minus the pore: redraw inside it
            const Pore& P = D.pores[S.pore];
            for (int k = 0; k < kMaxPosRedraw; ++k) {
                const double dx = pos.x - P.cx, dy = pos.y - P.cy;
                if (!(dx * dx + dy * dy < P.r2)) break;
                const int w = 4 * (B_POSRE + 2 * k);
                c[0] = g.u53(w);
                c[1] = g.u53(w + 1);
                c[2] = g.u53(w + 2);
                pos = add(S.o, mul(S.mat, v3(c[0], c[1], c[2])));
            }
        }
        dir = mul(S.rotG, draw_aniso(g, true));
        sign = dot(dir, S.gradT) < 0.0;
        sdom = e.index;
    }
    ph.alive = true;
    ph.sign = sign ? 1 : -1;
    ph.pos = pos;
    ph.dir = normalized(dir);
    ph.scatNext = 0.0;
    ph.nscat = 0;
    draw_scat_next(M, ph, g);
    return ph;
}

bool is_inside(const Domain& D, const Sdom& S, V3 p) {
    for (int b = S.pbeg; b < S.pend; ++b) {
        const Plane& P = D.planes[b];
        if (dot(P.n, p) + P.c < -S.eps) return false;
    }
    if (S.pore >= 0) {
        const Pore& Q = D.pores[S.pore];
        const double dx = p.x - Q.cx, dy = p.y - Q.cy;
        const double rr = Q.r - Q.eps;
        if (dx * dx + dy * dy < rr * rr) return false;
    }
    return true;
}

// subdomain.cpp:161-192 with phonon.cpp:95-105; returns the hit plane or -1.
// `dist` receives the distance moved.
int advect(const Domain& D, const Sdom& S, Phonon& ph, double& dist) {
    double minDistance = ph.scatNext;
    int hit = -1;
    for (int b = S.pbeg; b < S.pend; ++b) {
        const Plane& P = D.planes[b];
        double nd = dot(P.n, ph.dir);
        if (nd >= 0.0) continue;
        double t = -(P.c + dot(P.n, ph.pos)) / nd;
        if (t < minDistance) {
            minDistance = t;
            hit = b;
        }
    }
    if (S.pore >= 0) {   // entering root of the ray with the pore's cylinder
        const Pore& Q = D.pores[S.pore];
        const double dx = ph.pos.x - Q.cx, dy = ph.pos.y - Q.cy;
        const double bq = dx * ph.dir.x + dy * ph.dir.y;
        if (bq < 0.0) {
            const double aq = ph.dir.x * ph.dir.x + ph.dir.y * ph.dir.y;
            const double cq = dx * dx + dy * dy - Q.r2;
            const double disc = bq * bq - aq * cq;
            if (disc >= 0.0) {
                // a start on (or, by rounding, just inside) the wall heading
                // in meets it at once
                const double t = cq <= 0.0 ? 0.0 : cq / (std::sqrt(disc) - bq);
                if (t < minDistance) {
                    minDistance = t;
                    hit = -3;
                }
            }
        }
    }
    ph.scatNext -= minDistance;
    if (ph.scatNext < DBL_MIN) ph.scatNext = 0.0;
    set_pos(ph, add(ph.pos, scale(minDistance, ph.dir)));   // Phonon::move, phonon.cpp:95-105
    dist = minDistance;
    bool negDist = minDistance < -S.eps;
    bool outside = !is_inside(D, S, ph.pos);
    if (negDist || outside) {
        ph.alive = false;
        return -2;
    }
    return hit;
}

/* -------------------------------------------------------------- field */
struct Field {
    int64_t rows, cols;
    std::vector<double> data;  // column-major rows x cols
    void addcol(int64_t c, const double* amt, double f) {
        double* dst = &data[c * rows];
        for (int64_t r = 0; r < rows; ++r) dst[r] += amt[r] * f;
    }
    void addcol1(int64_t c, const double* amt) {
        double* dst = &data[c * rows];
        for (int64_t r = 0; r < rows; ++r) dst[r] += amt[r];
    }
};

V3 coord(const Sdom& S, V3 p) {
    V3 l = mul(S.inv, sub(p, S.o));
    return v3(double(S.div[0]) * l.x, double(S.div[1]) * l.y, double(S.div[2]) * l.z);
}

void coord2index(const Sdom& S, V3 c, int64_t idx[3]) {
    double cc[3] = {c.x, c.y, c.z};
    for (int k = 0; k < 3; ++k) {
        int64_t v = int64_t(std::floor(cc[k]));
        idx[k] = std::min(std::max<int64_t>(v, 0), S.max[k]);
    }
}

int64_t colOf(const Sdom& S, const int64_t idx[3]) {
    return S.colOff + idx[0] * S.stride[0] + idx[1] * S.stride[1] + idx[2] * S.stride[2];
}

// field.cpp:92-220
void accumulate(Field& F, const Sdom& S, V3 bpos, V3 epos, const double* amount) {
    int flag = S.accum;
    if (flag < -1) return;
    if (flag < 0) {
        F.addcol1(S.colOff, amount);
        return;
    }
    V3 bc = coord(S, bpos), ec = coord(S, epos);
    V3 dc = sub(ec, bc);
    int64_t bi[3], ei[3];
    coord2index(S, bc, bi);
    coord2index(S, ec, ei);
    double bcv[3] = {bc.x, bc.y, bc.z}, ecv[3] = {ec.x, ec.y, ec.z}, dcv[3] = {dc.x, dc.y, dc.z};
    const int64_t R = F.rows;
    if (flag < 3) {
        int d = flag;
        int64_t b = bi[d], e = ei[d];
        int64_t vb[3] = {0, 0, 0}, ve[3] = {0, 0, 0};
        vb[d] = b;
        ve[d] = e;
        if (b == e) {
            F.addcol1(colOf(S, vb), amount);
            return;
        }
        std::vector<double> cell(R);
        double ad = std::fabs(dcv[d]);
        for (int64_t r = 0; r < R; ++r) cell[r] = amount[r] / ad;
        int64_t pm;
        if (b < e) {
            F.addcol(colOf(S, vb), cell.data(), double(1 + b) - bcv[d]);
            F.addcol(colOf(S, ve), cell.data(), ecv[d] - double(e));
            pm = 1;
        } else {
            F.addcol(colOf(S, vb), cell.data(), bcv[d] - double(b));
            F.addcol(colOf(S, ve), cell.data(), double(1 + e) - ecv[d]);
            pm = -1;
        }
        int64_t nv[3] = {0, 0, 0};
        for (int64_t n = b + pm; n != e; n += pm) {
            nv[d] = n;
            F.addcol1(colOf(S, nv), cell.data());
        }
        return;
    }
    struct Step {
        int64_t s[3];
    };
    std::map<double, Step> borders;
    auto ins = borders.begin();
    for (int d = 0; d < 3; ++d) {
        if (std::fabs(dcv[d]) < DBL_MIN) continue;
        int64_t b = bi[d], e = ei[d];
        int64_t pm;
        if (b == e) continue;
        if (b < e) {
            b++;
            e++;
            pm = 1;
        } else {
            pm = -1;
        }
        ins = borders.begin();
        Step st{{0, 0, 0}};
        st.s[d] = pm;
        bool search = !borders.empty();
        for (int64_t n = b; n != e; n += pm) {
            double param = (double(n) - bcv[d]) / dcv[d];
            if (search) {
                auto found = borders.find(param);
                if (found != borders.end()) {
                    for (int k = 0; k < 3; ++k) found->second.s[k] += st.s[k];
                    ins = found;
                    continue;
                }
            }
            ins = borders.insert(ins, std::make_pair(param, st));
        }
    }
    borders.insert(ins, std::make_pair(1.0, Step{{0, 0, 0}}));
    int64_t index[3] = {bi[0], bi[1], bi[2]};
    double param = 0.0;
    for (auto it = borders.begin(); it != borders.end(); ++it) {
        F.addcol(colOf(S, index), amount, it->first - param);
        param = it->first;
        for (int k = 0; k < 3; ++k) index[k] += it->second.s[k];
    }
}

/* -------------------------------------------------------------- solve */
struct Counters {
    int64_t steps = 0, escapes = 0, absorbed = 0, truncated = 0, bundles = 0;
};

void record(double* trace, int64_t cap, uint32_t nstep, int hit, double dist, V3 apos,
            const Phonon& ph, int sdom, int fate) {
    if (!trace || int64_t(nstep) > cap) return;
    double* r = trace + (int64_t(nstep) - 1) * MCG_TRACE_WIDTH;
    r[0] = nstep;
    r[1] = hit;
    r[2] = dist;
    r[3] = apos.x;
    r[4] = apos.y;
    r[5] = apos.z;
    r[6] = ph.dir.x;
    r[7] = ph.dir.y;
    r[8] = ph.dir.z;
    r[9] = ph.scatNext;
    r[10] = ph.mode;
    r[11] = double(ph.nscat);
    r[12] = sdom;
    r[13] = fate;
    r[14] = ph.pos.x;
    r[15] = ph.pos.y;
    r[16] = ph.pos.z;
}

// The step loop of FieldProblem::solve (problem.cpp:401-437) for bundle n
// starting as ph in subdomain sdom.
void run_phonon(const Material& M, const Domain& D, const Problem& P, int64_t n, Phonon ph, int sdom,
                Field& F, Counters& C, mcg_counts* counts, double* trace, int64_t cap);

void run_bundle(const Material& M, const Domain& D, const Problem& P, int64_t n, Field& F,
                Counters& C, mcg_counts* counts, double* trace = nullptr, int64_t cap = 0) {
    int64_t emitIndex = std::upper_bound(P.emitCdf.begin(), P.emitCdf.end(), n) - P.emitCdf.begin();
    const Emitter& e = D.emitters[emitIndex];
    EventRng g0(P.seed, 0u, uint64_t(n));
    int sdom;
    Phonon ph = emit(M, D, e, g0, sdom);
    run_phonon(M, D, P, n, ph, sdom, F, C, counts, trace, cap);
}

void run_phonon(const Material& M, const Domain& D, const Problem& P, int64_t n, Phonon ph, int sdom,
                Field& F, Counters& C, mcg_counts* counts, double* trace, int64_t cap) {
    uint32_t nbdry = 0, nstep = 0, fate = MCG_MAXLOOP;
    double amt[64];
    std::vector<double> amtv;
    double* amount = amt;
    if (P.rows > 64) {
        amtv.resize(P.rows);
        amount = amtv.data();
    }
    int64_t maxloop = P.maxloop;
    for (int64_t i = 0; i < maxloop; ++i) {
        Phonon pre = ph;
        const Sdom& S = D.sdoms[sdom];
        double vel = M.modeVel[ph.mode];
        double dist;
        int bdry = advect(D, S, ph, dist);
        ++nstep;
        const V3 apos = ph.pos;
        if (!ph.alive) {
            C.escapes++;
            fate = MCG_ESCAPED;
            record(trace, cap, nstep, bdry, dist, apos, ph, sdom, fate);
            break;
        }
        // tally amount (problem.cpp:414, 473-637): dt = d / v
        double s = double(ph.sign);
        double dtime = dist / vel;
        V3 dpos = sub(ph.pos, pre.pos);
        switch (P.tally) {
            case MCG_TEMP: amount[0] = s * dtime; break;
            case MCG_FLUX:
                amount[0] = s * dpos.x;
                amount[1] = s * dpos.y;
                amount[2] = s * dpos.z;
                break;
            case MCG_MULTI:
                amount[0] = s * dtime;
                amount[1] = s * dpos.x;
                amount[2] = s * dpos.y;
                amount[3] = s * dpos.z;
                break;
            case MCG_CUMTEMP: {
                for (int64_t r = 0; r < P.rows; ++r) amount[r] = 0.0;
                int64_t idx = (pre.nscat + P.step - 1) / P.step;
                amount[idx] = s * dtime;
                break;
            }
            case MCG_CUMFLUX: {
                for (int64_t r = 0; r < P.rows; ++r) amount[r] = 0.0;
                int64_t idx = 3 * ((pre.nscat + P.step - 1) / P.step);
                amount[idx] = s * dpos.x;
                amount[idx + 1] = s * dpos.y;
                amount[idx + 2] = s * dpos.z;
                break;
            }
        }
        accumulate(F, S, pre.pos, ph.pos, amount);
        EventRng g(P.seed, uint32_t(i + 1), uint64_t(n));
        if (bdry == -3) {
            ++nbdry;
            const Pore& Q = D.pores[S.pore];
            const V3 nn = normalized(v3((ph.pos.x - Q.cx) / Q.r, (ph.pos.y - Q.cy) / Q.r, 0.0));
            int kind = Q.kind;
            if (kind == MCG_PARTIAL) kind = g.word(W_SPEC) < Q.thr ? MCG_SPEC : MCG_DIFF;
            if (kind == MCG_SPEC) {
                ph.dir = normalized(mul(refl_matrix(nn), ph.dir));
            } else {
                ph.dir = normalized(mul(rot_unit(nn), draw_aniso(g, false)));
                ph.nscat++;
            }
        } else if (bdry >= 0) {
            ++nbdry;
            const Plane& B = D.planes[bdry];
            int kind = B.kind;
            if (kind == MCG_PARTIAL) kind = g.word(W_SPEC) < B.specThr ? MCG_SPEC : MCG_DIFF;
            switch (kind) {
                case MCG_SPEC:  // boundary.cpp:283-287
                    ph.dir = normalized(mul(B.refl, ph.dir));
                    sdom = B.sdom;
                    break;
                case MCG_DIFF:  // boundary.cpp:308-312
                    ph.dir = normalized(mul(B.rot, draw_aniso(g, false)));
                    ph.nscat++;
                    sdom = B.sdom;
                    break;
                case MCG_ISOT:  // boundary.cpp:456-460
                    ph.alive = false;
                    fate = MCG_ABSORBED;
                    sdom = B.sdom;
                    break;
                case MCG_PERI: {  // boundary.cpp:517-522
                    set_pos(ph, add(mul(B.periRot, ph.pos), B.periT));
                    ph.dir = normalized(mul(B.periRot, ph.dir));
                    sdom = D.planes[B.pair].sdom;
                    break;
                }
                case MCG_INTER: {  // boundary.cpp:349-359
                    int target = -1;
                    if (B.pairs.size() == 1) {
                        target = B.pairs[0];
                    } else {
                        for (int q : B.pairs) {
                            const Sdom& T = D.sdoms[D.planes[q].sdom];
                            if (is_inside(D, T, ph.pos)) {
                                target = q;
                                break;
                            }
                        }
                    }
                    if (target < 0) {
                        ph.alive = false;
                        C.escapes++;
                        fate = MCG_ESCAPED;
                        record(trace, cap, nstep, bdry, dist, apos, ph, sdom, fate);
                        goto done;
                    }
                    sdom = D.planes[target].sdom;
                    break;
                }
            }
        } else {
            // material.cpp:226-231
            ph.mode = alias_draw(M.scatAlias, g.word(W_MODE));
            ph.dir = normalized(draw_iso(g));
            ph.nscat++;
            draw_scat_next(M, ph, g);
        }
        int f = -1;
        if (!ph.alive) f = MCG_ABSORBED;
        else if (ph.nscat >= P.maxscat) f = MCG_MAXSCAT;
        else if (i + 1 >= maxloop) f = MCG_MAXLOOP;
        record(trace, cap, nstep, bdry, dist, apos, ph, sdom, f);
        if (!ph.alive) break;
        if (ph.nscat >= P.maxscat) {
            fate = MCG_MAXSCAT;
            break;
        }
    }
done:
    C.steps += nstep;
    C.bundles++;
    if (fate == MCG_ABSORBED) C.absorbed++;
    if (fate == MCG_MAXSCAT || fate == MCG_MAXLOOP) C.truncated++;
    if (counts) {
        int64_t k = n - P.first;
        if (counts->nscat) counts->nscat[k] = uint32_t(ph.nscat);
        if (counts->nbdry) counts->nbdry[k] = nbdry;
        if (counts->nstep) counts->nstep[k] = nstep;
        if (counts->fate) counts->fate[k] = fate;
    }
}

// postProc + normalisation (problem.cpp:439-444 and the kinds' postProc)
void normalise(const Material& M, const Domain& D, const Problem& P, const Field& F,
               std::vector<double>& out) {
    const int64_t R = F.rows, Cn = F.cols;
    std::vector<double> sol(F.data);
    switch (P.tally) {
        case MCG_TEMP:
            for (double& v : sol) v = v / M.energySum;
            break;
        case MCG_FLUX: break;
        case MCG_MULTI:
            for (int64_t c = 0; c < Cn; ++c) sol[c * R] = sol[c * R] / M.energySum;
            break;
        case MCG_CUMTEMP:
            for (int64_t c = 0; c < Cn; ++c) {
                for (int64_t i = 0; i < P.size; ++i) sol[c * R + i + 1] += sol[c * R + i];
                for (int64_t r = 0; r < R; ++r) sol[c * R + r] = sol[c * R + r] / M.energySum;
            }
            break;
        case MCG_CUMFLUX:
            for (int64_t c = 0; c < Cn; ++c)
                for (int64_t i = 0; i < P.size; ++i)
                    for (int k = 0; k < 3; ++k)
                        sol[c * R + 3 * (i + 1) + k] += sol[c * R + 3 * i + k];
            break;
    }
    out.assign(R * Cn, 0.0);
    for (int64_t c = 0; c < Cn; ++c)
        // a cell with no material (wholly inside a pore) reports zero
        for (int64_t r = 0; r < R; ++r)
            out[c * R + r] = D.cellVol[c] > 0.0 ? P.power * (sol[c * R + r] / D.cellVol[c]) : 0.0;
}

struct Built {
    Material M;
    Domain D;
    Problem P;
};

int build_all(const mcg_material_spec* mat, const mcg_domain_spec* dom,
              const mcg_problem_spec* prob, Built& B, bool trace_only = false) {
    int rc = build_material(mat, B.M);
    if (rc) return rc;
    rc = build_domain(dom, B.D);
    if (rc) return rc;
    return build_problem(prob, B.M, B.D, B.P, trace_only);
}

}  // namespace

extern "C" {

const char* oracle_last_error(void) { return g_err.c_str(); }

void oracle_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    philox(ctr, key[0], key[1], out);
}

int oracle_alias(const double* pdf, int64_t n, uint32_t* thr, uint32_t* alias) {
    Alias a;
    int rc = build_alias(pdf, n, a);
    if (rc) return rc;
    std::memcpy(thr, a.thr.data(), sizeof(uint32_t) * n);
    std::memcpy(alias, a.alias.data(), sizeof(uint32_t) * n);
    return MCG_OK;
}

int oracle_material(const mcg_material_spec* mat, mcg_material_tables* out) {
    Material M;
    int rc = build_material(mat, M);
    if (rc) return rc;
    size_t n = size_t(M.nw * M.np);
    if (out->tau) std::memcpy(out->tau, M.tau.data(), n * sizeof(double));
    if (out->vel) std::memcpy(out->vel, M.vel.data(), n * sizeof(double));
    if (out->energy_pdf) std::memcpy(out->energy_pdf, M.energy.data(), n * sizeof(double));
    if (out->flux_pdf) std::memcpy(out->flux_pdf, M.flux.data(), n * sizeof(double));
    if (out->scat_pdf) std::memcpy(out->scat_pdf, M.scat.data(), n * sizeof(double));
    out->energy_sum = M.energySum;
    out->flux_sum = M.fluxSum;
    out->scat_sum = M.scatSum;
    out->cond = M.k;
    out->nmodes = int64_t(M.modeCell.size());
    return MCG_OK;
}

int oracle_plan(const mcg_material_spec* mat, const mcg_domain_spec* dom,
                const mcg_problem_spec* prob, mcg_problem_info* info, double* w, int64_t* cnt,
                int64_t cap) {
    Built B;
    int rc = build_all(mat, dom, prob, B);
    if (rc) return rc;
    info->nemit = B.P.nemit;
    info->rows = B.P.rows;
    info->cols = B.D.cols;
    info->maxloop = B.P.maxloop;
    info->nemitter = int64_t(B.D.emitters.size());
    info->power = B.P.power;
    info->weight_sum = B.P.weightSum;
    for (int64_t i = 0; i < info->nemitter && i < cap; ++i) {
        if (w) w[i] = B.D.emitters[i].weight;
        if (cnt) cnt[i] = B.P.emitPdf[i];
    }
    return MCG_OK;
}

// Cell volumes of the field columns (field.cpp:53-77 with CellVolF).
int oracle_cell_volumes(const mcg_domain_spec* dom, double* out, int64_t cap, int64_t* ncols) {
    Domain D;
    int rc = build_domain(dom, D);
    if (rc) return rc;
    *ncols = D.cols;
    for (int64_t i = 0; i < D.cols && i < cap; ++i) out[i] = D.cellVol[i];
    return MCG_OK;
}

// One Field::accumulate call on subdomain `sdom` for a rows-long amount;
// `field` is rows x cols column-major and is added to.
int oracle_accumulate(const mcg_domain_spec* dom, int32_t sdom, const double bpos[3],
                      const double epos[3], int64_t rows, const double* amount, double* field,
                      int64_t field_len) {
    Domain D;
    int rc = build_domain(dom, D);
    if (rc) return rc;
    if (sdom < 0 || sdom >= int(D.sdoms.size())) return fail(MCG_EINVAL, "bad subdomain");
    if (field_len < rows * D.cols) return fail(MCG_ERANGE, "field buffer too small");
    Field F{rows, D.cols, std::vector<double>(field, field + rows * D.cols)};
    accumulate(F, D.sdoms[sdom], v3(bpos[0], bpos[1], bpos[2]), v3(epos[0], epos[1], epos[2]),
               amount);
    std::memcpy(field, F.data.data(), sizeof(double) * rows * D.cols);
    return MCG_OK;
}

// Direction samplers for statistical tests: kind 0 = drawIso, 1 = drawAniso
// (Lambertian), 2 = drawAniso bidirectional.  out is n x 3.
void oracle_sample_dirs(int32_t kind, uint64_t seed, int64_t n, double* out) {
    for (int64_t i = 0; i < n; ++i) {
        EventRng g(seed, 1u, uint64_t(i));
        V3 d = kind == 0 ? draw_iso(g) : draw_aniso(g, kind == 2);
        out[3 * i] = d.x;
        out[3 * i + 1] = d.y;
        out[3 * i + 2] = d.z;
    }
}

// Emission draws for statistical tests: n bundles of the problem, each
// giving (x, y, z, dx, dy, dz, sign, mode, scatNext, emitter).
int oracle_sample_emission(const mcg_material_spec* mat, const mcg_domain_spec* dom,
                           const mcg_problem_spec* prob, int64_t n, double* out) {
    Built B;
    int rc = build_all(mat, dom, prob, B);
    if (rc) return rc;
    for (int64_t k = 0; k < n; ++k) {
        int64_t b = k % B.P.nemit;
        int64_t ei = std::upper_bound(B.P.emitCdf.begin(), B.P.emitCdf.end(), b) -
                     B.P.emitCdf.begin();
        EventRng g(B.P.seed, 0u, uint64_t(k));
        int sd;
        Phonon ph = emit(B.M, B.D, B.D.emitters[ei], g, sd);
        double* o = out + 10 * k;
        o[0] = ph.pos.x;
        o[1] = ph.pos.y;
        o[2] = ph.pos.z;
        o[3] = ph.dir.x;
        o[4] = ph.dir.y;
        o[5] = ph.dir.z;
        o[6] = ph.sign;
        o[7] = double(B.M.modeCell[ph.mode]);
        o[8] = ph.scatNext;
        o[9] = double(ei);
    }
    return MCG_OK;
}

// Rectangle-disk overlap areas, element-wise: q holds 7 doubles per query
// (x0, x1, y0, y1, cx, cy, r).
void oracle_disk_area(const double* q, int64_t n, double* out) {
    for (int64_t i = 0; i < n; ++i) {
        const double* a = q + 7 * i;
        out[i] = mcg_disk::rect_in_disk(a[0], a[1], a[2], a[3], a[4], a[5], a[6]);
    }
}

// The portable math functions, element-wise (for accuracy tests).
void oracle_math(int32_t fn, const double* x, int64_t n, double* out) {
    for (int64_t i = 0; i < n; ++i) {
        if (fn == 0) {
            out[i] = mcg_math::log(x[i]);
        } else {
            double s, c;
            mcg_math::sincospi(x[i], &s, &c);
            out[i] = fn == 1 ? s : c;
        }
    }
}

// Step-by-step record of one bundle (see mcg_trace in include/mcg.h).
int oracle_trace(const mcg_material_spec* mat, const mcg_domain_spec* dom,
                 const mcg_problem_spec* prob, int64_t bundle, double* rec, int64_t cap,
                 int64_t* nrec) {
    Built B;
    int rc = build_all(mat, dom, prob, B);
    if (rc) return rc;
    if (bundle < 0 || bundle >= B.P.nemit) return fail(MCG_ERANGE, "bundle outside [0, nemit)");
    Field F{B.P.rows, B.D.cols, std::vector<double>(B.P.rows * B.D.cols, 0.0)};
    Counters C;
    uint32_t ns, nb, nt, fa;
    mcg_counts counts{&ns, &nb, &nt, &fa};
    Problem P1 = B.P;
    P1.first = bundle;
    run_bundle(B.M, B.D, P1, bundle, F, C, &counts, rec, cap);
    *nrec = std::min<int64_t>(nt, cap);
    return MCG_OK;
}

// TrajProblem::solve from a given start (problem.cpp:226-299, pos_ set):
// the mode from the scattering distribution unless given (drawScatProp),
// an isotropic direction unless given (drawIso), the first free path, all
// from event 0 of bundle prob->first_bundle; the start cell by
// Domain::locate (domain.cpp:58-66).  The domain needs no source.  With
// track set the trajectory's points are collected as TrkPhonon does: the
// start, then every position the bundle takes (phonon.cpp:133-157).
static int trace_from(const mcg_material_spec* mat, const mcg_domain_spec* dom,
                      const mcg_problem_spec* prob, int64_t prop_cell, const double pos[3], const double* dir, double* rec, int64_t cap,
               int64_t* nrec, std::vector<V3>* track) {
    if (!prob || !pos) return fail(MCG_EINVAL, "NULL argument");
    if (prob->first_bundle < 0) return fail(MCG_ERANGE, "bundle id must be non-negative");
    mcg_problem_spec one = *prob;
    one.first_bundle = 0;
    one.bundle_count = 1;
    Built B;
    int rc = build_all(mat, dom, &one, B, true);
    if (rc) return rc;
    B.P.first = prob->first_bundle;   // the stream's id; it need not be an emitted bundle
    int mode = -1;
    if (prop_cell >= 0) {
        for (size_t k = 0; k < B.M.modeCell.size(); ++k)
            if (B.M.modeCell[k] == prop_cell) mode = int(k);
        if (mode < 0) return fail(MCG_EINVAL, "prop has no energy pdf");
    }
    const V3 p = v3(pos[0], pos[1], pos[2]);
    int sdom = -1;
    for (size_t si = 0; si < B.D.sdoms.size() && sdom < 0; ++si)
        if (is_inside(B.D, B.D.sdoms[si], p)) sdom = int(si);
    if (sdom < 0) return fail(MCG_EINVAL, "Position not inside domain");
    if (dir && !(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2] > 0.0))
        return fail(MCG_EINVAL, "direction must be nonzero");
    const int64_t n = B.P.first;
    EventRng g(B.P.seed, 0u, uint64_t(n));
    Phonon ph{};
    ph.mode = mode >= 0 ? mode : alias_draw(B.M.scatAlias, g.word(W_MODE));
    ph.dir = normalized(dir ? v3(dir[0], dir[1], dir[2]) : draw_iso(g));
    ph.pos = p;
    ph.alive = true;
    ph.sign = 1;
    ph.scatNext = 0.0;
    ph.nscat = 0;
    ph.track = track;
    if (track) track->push_back(p);
    draw_scat_next(B.M, ph, g);
    Field F{B.P.rows, B.D.cols, std::vector<double>(B.P.rows * B.D.cols, 0.0)};
    Counters C;
    uint32_t ns, nb, nt, fa;
    mcg_counts counts{&ns, &nb, &nt, &fa};
    run_phonon(B.M, B.D, B.P, n, ph, sdom, F, C, &counts, rec, cap);
    if (nrec) *nrec = std::min<int64_t>(nt, cap);
    return MCG_OK;
}

int oracle_trace_from(const mcg_material_spec* mat, const mcg_domain_spec* dom,
                      const mcg_problem_spec* prob, int64_t prop_cell, const double pos[3],
                      const double* dir, double* rec, int64_t cap, int64_t* nrec) {
    if (!rec || !nrec || cap <= 0) return fail(MCG_EINVAL, "NULL argument");
    return trace_from(mat, dom, prob, prop_cell, pos, dir, rec, cap, nrec, nullptr);
}

// The 3 x n points TrajProblem::solve returns (phn.trajectory(),
// phonon.cpp:159-169), column-major into pts (cap points at most); *npts
// receives the full count.
int oracle_trajectory_from(const mcg_material_spec* mat, const mcg_domain_spec* dom,
                           const mcg_problem_spec* prob, int64_t prop_cell, const double pos[3],
                           const double* dir, double* pts, int64_t cap, int64_t* npts) {
    if (!pts || !npts || cap <= 0) return fail(MCG_EINVAL, "NULL argument");
    std::vector<V3> track;
    int rc = trace_from(mat, dom, prob, prop_cell, pos, dir, nullptr, 0, nullptr, &track);
    if (rc) return rc;
    *npts = int64_t(track.size());
    for (int64_t i = 0; i < *npts && i < cap; ++i) {
        pts[3 * i] = track[size_t(i)].x;
        pts[3 * i + 1] = track[size_t(i)].y;
        pts[3 * i + 2] = track[size_t(i)].z;
    }
    return MCG_OK;
}

// The whole field solve: solveField (main.cpp:146-179) around
// FieldProblem::solve (problem.cpp:370-445).  nthreads <= 0 uses OpenMP's
// default.  Each thread runs a static share of the bundle range into its own
// field; the normalised partials are summed in thread order.
int oracle_solve(const mcg_material_spec* mat, const mcg_domain_spec* dom,
                 const mcg_problem_spec* prob, int32_t nthreads, double* out, int64_t out_len,
                 mcg_counts* counts, mcg_stats* stats) {
    auto t0 = std::chrono::steady_clock::now();
    Built B;
    int rc = build_all(mat, dom, prob, B);
    if (rc) return rc;
    const int64_t R = B.P.rows, Cn = B.D.cols;
    if (out_len < R * Cn) return fail(MCG_ERANGE, "output buffer too small");
    int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
    std::vector<std::vector<double>> partial(nt);
    std::vector<Counters> cnt(nt);
    auto t1 = std::chrono::steady_clock::now();
#pragma omp parallel num_threads(nt)
    {
        int t = omp_get_thread_num();
        Field F{R, Cn, std::vector<double>(R * Cn, 0.0)};
#pragma omp for schedule(static)
        for (int64_t n = B.P.first; n < B.P.first + B.P.count; ++n)
            run_bundle(B.M, B.D, B.P, n, F, cnt[t], counts);
        normalise(B.M, B.D, B.P, F, partial[t]);
    }
    auto t2 = std::chrono::steady_clock::now();
    for (int64_t i = 0; i < R * Cn; ++i) out[i] = 0.0;
    Counters tot;
    for (int t = 0; t < nt; ++t) {
        for (int64_t i = 0; i < R * Cn; ++i) out[i] += partial[t][i];
        tot.steps += cnt[t].steps;
        tot.escapes += cnt[t].escapes;
        tot.absorbed += cnt[t].absorbed;
        tot.truncated += cnt[t].truncated;
        tot.bundles += cnt[t].bundles;
    }
    if (stats) {
        stats->bundles = tot.bundles;
        stats->steps = tot.steps;
        stats->escapes = tot.escapes;
        stats->absorbed = tot.absorbed;
        stats->truncated = tot.truncated;
        stats->kernel_ms = std::chrono::duration<double, std::milli>(t2 - t1).count();
        stats->wall_ms = std::chrono::duration<double, std::milli>(t2 - t0).count();
    }
    return MCG_OK;
}

}  // extern "C"
