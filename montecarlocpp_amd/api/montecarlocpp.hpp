// montecarlocpp.hpp -- the reference's C++ solve API, routed to the GPU.
//
// Header-only C++ mirror of the classes a montecarlocpp caller uses to set up
// and run a field solve: Material (material.h:22-70), the boundary family
// (boundary.h:33-281), the cell templates (subdomain.h:35-352), Domain with
// BulkDomain / FilmDomain (domain.h:37-137), Field (field.h:26-49),
// Statistics (field.h:51-66), Problem / FieldProblem and the five tally kinds
// (problem.h:65-239), Progress (problem.h:42-60) and solveField / calcStats
// (main.cpp:146-210).  Names, constructor arguments and their meaning follow
// the reference, so a deck written against it compiles against this header;
// FieldProblem::solve hands the whole solve to the MI355X engine through the
// C ABI of include/mcg.h (one mcg_solve per solve, not per OpenMP thread).
//
// Eigen and Boost are not used; the small Vector3d / Matrix3d / Vector3l /
// ArrayXXd types below cover what the API passes around.  Errors that the
// reference asserts on are thrown as mcg::Error with the same message.
#pragma once

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <fstream>
#include <initializer_list>
#include <iostream>
#include <map>
#include <memory>
#include <random>
#include <tuple>
#include <utility>
#include <sstream>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "mcg.h"

namespace mcg {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

inline void check(int rc) {
    if (rc != MCG_OK) throw Error(rc, mcg_last_error());
}

}  // namespace mcg

// --------------------------------------------------------------- small algebra
struct Vector3d {
    double v[3] = {0, 0, 0};
    Vector3d() = default;
    Vector3d(double x, double y, double z) : v{x, y, z} {}
    double& operator()(int i) { return v[i]; }
    double operator()(int i) const { return v[i]; }
    static Vector3d Zero() { return Vector3d(); }
    static Vector3d Unit(int i) { Vector3d r; r.v[i] = 1.0; return r; }
    Vector3d operator+(const Vector3d& o) const { return {v[0] + o.v[0], v[1] + o.v[1], v[2] + o.v[2]}; }
    Vector3d operator-(const Vector3d& o) const { return {v[0] - o.v[0], v[1] - o.v[1], v[2] - o.v[2]}; }
    Vector3d operator*(double s) const { return {v[0] * s, v[1] * s, v[2] * s}; }
};
inline Vector3d operator*(double s, const Vector3d& a) { return a * s; }

struct Vector3l {
    long v[3] = {0, 0, 0};
    Vector3l() = default;
    Vector3l(long x, long y, long z) : v{x, y, z} {}
    long& operator()(int i) { return v[i]; }
    long operator()(int i) const { return v[i]; }
};

// 3x3, column-major like Eigen's Matrix3d
struct Matrix3d {
    double m[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    double& operator()(int r, int c) { return m[3 * c + r]; }
    double operator()(int r, int c) const { return m[3 * c + r]; }
    static Matrix3d Identity() { Matrix3d a; a.m[0] = a.m[4] = a.m[8] = 1.0; return a; }
    static Matrix3d Zero() { return Matrix3d(); }
    static Matrix3d Diagonal(const Vector3d& d) {
        Matrix3d a;
        a.m[0] = d(0); a.m[4] = d(1); a.m[8] = d(2);
        return a;
    }
    Vector3d col(int c) const { return {m[3 * c], m[3 * c + 1], m[3 * c + 2]}; }
};

using VectorXd = std::vector<double>;

// 3 x N, column-major like Eigen's Matrix3Xd
class Matrix3Xd {
    std::vector<double> d_;

public:
    Matrix3Xd() = default;
    Matrix3Xd(long rows, long cols) : d_(size_t(3 * cols), 0.0) {
        if (rows != 3) throw mcg::Error(MCG_EINVAL, "Matrix3Xd has three rows");
    }
    long cols() const { return long(d_.size() / 3); }
    double& operator()(int r, int c) { return d_[size_t(3 * c + r)]; }
    double operator()(int r, int c) const { return d_[size_t(3 * c + r)]; }
    Vector3d col(int c) const { return {d_[3 * c], d_[3 * c + 1], d_[3 * c + 2]}; }
    const double* data() const { return d_.data(); }
};

// Column-major rows x cols array (Eigen's ArrayXXd layout).
class ArrayXXd {
    long rows_ = 0, cols_ = 0;
    std::vector<double> d_;

public:
    ArrayXXd() = default;
    ArrayXXd(long r, long c) : rows_(r), cols_(c), d_(size_t(r * c), 0.0) {}
    static ArrayXXd Zero(long r, long c) { return ArrayXXd(r, c); }
    long rows() const { return rows_; }
    long cols() const { return cols_; }
    long size() const { return rows_ * cols_; }
    double* data() { return d_.data(); }
    const double* data() const { return d_.data(); }
    double& operator()(long r, long c) { return d_[size_t(c * rows_ + r)]; }
    double operator()(long r, long c) const { return d_[size_t(c * rows_ + r)]; }
    ArrayXXd& operator+=(const ArrayXXd& o) {
        for (size_t i = 0; i < d_.size(); ++i) d_[i] += o.d_[i];
        return *this;
    }
    ArrayXXd operator-(const ArrayXXd& o) const { ArrayXXd r(*this); for (size_t i = 0; i < d_.size(); ++i) r.d_[i] -= o.d_[i]; return r; }
    ArrayXXd operator*(const ArrayXXd& o) const { ArrayXXd r(*this); for (size_t i = 0; i < d_.size(); ++i) r.d_[i] *= o.d_[i]; return r; }
    ArrayXXd operator/(double s) const { ArrayXXd r(*this); for (double& x : r.d_) x /= s; return r; }
    ArrayXXd operator+(const ArrayXXd& o) const { ArrayXXd r(*this); r += o; return r; }
    ArrayXXd sqrt() const { ArrayXXd r(*this); for (double& x : r.d_) x = std::sqrt(x); return r; }
};

// ------------------------------------------------------------------ Rng
// The reference's Rng is a Boost mt19937 (random.h:22) seeded per thread.
// Here it only carries the seed of the solve's Philox key: the engine derives
// every bundle's stream from (seed, bundle id).
class Rng {
    std::mt19937_64 g_;

public:
    explicit Rng(uint64_t seed = 5489u) : g_(seed) {}
    uint64_t operator()() { return g_(); }
};

// --------------------------------------------------------------- Material
// Material(disp, relax, T): the reference's two text formats
// (material.cpp:87-114), parsed here; the tables are built by the engine.
class Material {
    long nw_ = 0, np_ = 0;
    double T_ = 0;
    std::vector<double> disp_, relax_;
    mcg_material_tables tab_{};
    std::string dispName_, relaxName_;

    void compute() {
        mcg_material_spec s = spec();
        mcg::check(mcg_material_compute(&s, &tab_));
    }

public:
    Material() = default;
    Material(const std::string& disp, const std::string& relax, double temp)
        : T_(temp), dispName_(disp), relaxName_(relax) {
        std::ifstream df(disp.c_str());
        if (!df) throw mcg::Error(MCG_EINVAL, "Error opening dispersion file");
        df >> nw_ >> np_;
        if (!(nw_ > 0 && np_ > 0)) throw mcg::Error(MCG_EINVAL, "Invalid dispersion file");
        disp_.resize(size_t(nw_ * (2 + 2 * np_)));
        for (double& x : disp_)
            if (!(df >> x)) throw mcg::Error(MCG_EINVAL, "Array extraction failed");
        std::ifstream rf(relax.c_str());
        if (!rf) throw mcg::Error(MCG_EINVAL, "Error opening relaxation time file");
        relax_.resize(size_t(8 * np_));
        for (double& x : relax_)
            if (!(rf >> x)) throw mcg::Error(MCG_EINVAL, "Array extraction failed");
        compute();
    }
    Material(long nw, long np, std::vector<double> disp, std::vector<double> relax, double temp)
        : nw_(nw), np_(np), T_(temp), disp_(std::move(disp)), relax_(std::move(relax)) {
        compute();
    }
    mcg_material_spec spec() const {
        return mcg_material_spec{nw_, np_, disp_.data(), relax_.data(), T_};
    }
    double temp() const { return T_; }
    long nw() const { return nw_; }
    long np() const { return np_; }
    double cond() const { return tab_.cond; }
    double energySum() const { return tab_.energy_sum; }
    double fluxSum() const { return tab_.flux_sum; }
    double scatSum() const { return tab_.scat_sum; }
    friend std::ostream& operator<<(std::ostream& os, const Material& m) {
        return os << "Material " << &m << "\n  disp:  " << m.dispName_ << "\n  relax: "
                  << m.relaxName_ << "\n  temp:  " << m.T_;
    }
};

// --------------------------------------------------------------- Phonon
// The bundle state (phonon.h:20-71); on the GPU it lives in registers.  This
// host type is what a trajectory record describes.
class Phonon {
public:
    class Prop {
        long w_ = 0, p_ = 0;

    public:
        Prop() = default;
        Prop(long w, long p) : w_(w), p_(p) {}
        long w() const { return w_; }
        long p() const { return p_; }
    };
    bool alive = true;
    int sign = 1;
    Prop prop;
    Vector3d pos, dir;
    double scatNext = 0;
    long nscat = 0;
};

// ------------------------------------------------------------- boundaries
class Subdomain;

class Boundary {
public:
    class Shape {
    public:
        virtual ~Shape() = default;
    };
    virtual ~Boundary() = default;
    virtual int kind() const = 0;
    virtual std::string type() const = 0;
    double temp() const { return T_; }
    double specularity() const { return p_; }

protected:
    double T_ = 0.0;
    double p_ = 0.0;
    friend class Subdomain;
};

class Parallelogram : public Boundary::Shape {
    Vector3d i_, j_;

public:
    Parallelogram() = default;
    Parallelogram(const Vector3d& i, const Vector3d& j) : i_(i), j_(j) {}
};
class Triangle : public Boundary::Shape {
    Vector3d i_, j_;

public:
    Triangle() = default;
    Triangle(const Vector3d& i, const Vector3d& j) : i_(i), j_(j) {}
};
// Polygon<N> (boundary.h:112-130): the N - 1 vertices after the first,
// relative to it; the engine checks the fan when the scene is built.
template <int N>
class Polygon : public Boundary::Shape {
    static_assert(N > 3, "Polygon must have more than 3 sides");
    Matrix3Xd verts_;

public:
    Polygon() = default;
    explicit Polygon(const Matrix3Xd& v) : verts_(v) {
        if (v.cols() != N - 1) throw mcg::Error(MCG_EINVAL, "Incorrect number of vertices");
    }
};

// Non-emitting faces: a temperature is an error (boundary.cpp:272-326).
template <int K>
class PlainBoundary : public Boundary {
public:
    PlainBoundary() = default;
    template <typename S>
    PlainBoundary(const Vector3d&, const S&, double T) {
        if (T != 0.0) throw mcg::Error(MCG_EINVAL, "Cannot specify temperature");
    }
    int kind() const override { return K; }
    std::string type() const override {
        return K == MCG_SPEC ? "Spec" : K == MCG_DIFF ? "Diff" : "Inter";
    }
};
using SpecBoundary = PlainBoundary<MCG_SPEC>;
using DiffBoundary = PlainBoundary<MCG_DIFF>;

// Order of makePair calls: a face's Inter partners are tried in this order
// (boundary.cpp:349-359), so the flattened pair list keeps it.
inline long nextPairSeq() {
    static long seq = 0;
    return ++seq;
}

class InterBoundary : public PlainBoundary<MCG_INTER> {
public:
    using PlainBoundary<MCG_INTER>::PlainBoundary;
    std::vector<InterBoundary*> pairs;      // makePair partners added by this face
    std::vector<long> pairSeq;              // their makePair call order
};

// Extension: specular with probability p, else diffuse.
class PartialBoundary : public Boundary {
public:
    PartialBoundary() = default;
    template <typename S>
    PartialBoundary(const Vector3d&, const S&, double T) {
        if (T != 0.0) throw mcg::Error(MCG_EINVAL, "Cannot specify temperature");
    }
    void specularity(double p) { p_ = p; }
    double specularity() const { return p_; }
    int kind() const override { return MCG_PARTIAL; }
    std::string type() const override { return "Partial"; }
};

template <typename S>
class IsotBoundary : public Boundary {
public:
    IsotBoundary() = default;
    IsotBoundary(const Vector3d&, const S&, double T) { T_ = T; }
    int kind() const override { return MCG_ISOT; }
    std::string type() const override { return "Isot"; }
};

// The periodic map of a makePair call, kept on its first face.
class PeriBase : public Boundary {
public:
    int kind() const override { return MCG_PERI; }
    std::string type() const override { return "Peri"; }
    const PeriBase* pair = nullptr;
    long pairSeq = 0;
    Vector3d transl;
    Matrix3d rot = Matrix3d::Identity();
};

template <typename S>
class PeriBoundary : public PeriBase {
public:
    PeriBoundary() = default;
    PeriBoundary(const Vector3d&, const S&, double T) { T_ = T; }
};

template <typename S>
void makePair(PeriBoundary<S>& b1, PeriBoundary<S>& b2, const Vector3d& t,
              const Matrix3d& r = Matrix3d::Identity()) {
    if (b1.pair || b2.pair) throw mcg::Error(MCG_EINVAL, "Boundary already paired");
    b1.pair = &b2;
    b1.pairSeq = nextPairSeq();
    b1.transl = t;
    b1.rot = r;
}
inline void makePair(InterBoundary& b1, InterBoundary& b2) {
    b1.pairs.push_back(&b2);
    b1.pairSeq.push_back(nextPairSeq());
}

// --------------------------------------------------------------- cells
class Subdomain {
public:
    using Pointers = std::vector<const Subdomain*>;
    virtual ~Subdomain() = default;
    virtual std::vector<const Boundary*> bdryPtrs() const = 0;
    mcg_subdomain_spec spec() const {
        mcg_subdomain_spec s{};
        s.cell = cell_;
        for (int k = 0; k < 3; ++k) {
            s.origin[k] = o_(k);
            s.div[k] = div_(k);
            s.grad_t[k] = gradT_(k);
        }
        if (matx_.cols() > 0)
            for (long i = 0; i < 3 * matx_.cols(); ++i) s.mat[i] = matx_.data()[i];
        else
            for (int i = 0; i < 9; ++i) s.mat[i] = mat_.m[i];
        std::vector<const Boundary*> b = bdryPtrs();
        s.nface = int32_t(b.size());
        for (size_t f = 0; f < b.size(); ++f) {
            s.face[f].kind = b[f]->kind();
            s.face[f].temp = b[f]->T_;
            s.face[f].specularity = b[f]->p_;
        }
        return s;
    }

protected:
    int cell_ = MCG_PARALLELEPIPED;
    Vector3d o_, gradT_;
    Matrix3d mat_;
    Matrix3Xd matx_;   // Prism / Pyramid: the N columns of their mat
    Vector3l div_;
};
using EmitSubdomain = Subdomain;

// Parallelepiped<Bac, Lef, Bot, Fro, Rig, Top>(o, mat, div, gradT, T)
// (subdomain.h:128-198): faces at o (Bac, Lef, Bot) and o + mat.col(k).
template <typename Bac, typename Lef, typename Bot, typename Fro = Bac, typename Rig = Lef,
          typename Top = Bot>
class Parallelepiped : public Subdomain {
    std::tuple<Bac, Lef, Bot, Fro, Rig, Top> b_;

public:
    Parallelepiped(const Vector3d& o, const Matrix3d& mat, const Vector3l& div,
                   const Vector3d& gradT = Vector3d::Zero(),
                   const std::vector<double>& T = std::vector<double>(6, 0.0))
        : b_(Bac(o, Parallelogram(mat.col(1), mat.col(2)), T.at(0)),
             Lef(o, Parallelogram(mat.col(2), mat.col(0)), T.at(1)),
             Bot(o, Parallelogram(mat.col(0), mat.col(1)), T.at(2)),
             Fro(o + mat.col(0), Parallelogram(mat.col(2), mat.col(1)), T.at(3)),
             Rig(o + mat.col(1), Parallelogram(mat.col(0), mat.col(2)), T.at(4)),
             Top(o + mat.col(2), Parallelogram(mat.col(1), mat.col(0)), T.at(5))) {
        if (T.size() != 6) throw mcg::Error(MCG_EINVAL, "Incorrect number of temperatures");
        cell_ = MCG_PARALLELEPIPED;
        o_ = o;
        mat_ = mat;
        div_ = div;
        gradT_ = gradT;
    }
    Parallelepiped(const Parallelepiped&) = delete;
    template <int I>
    typename std::tuple_element<I, std::tuple<Bac, Lef, Bot, Fro, Rig, Top>>::type& bdry() {
        return std::get<I>(b_);
    }
    std::vector<const Boundary*> bdryPtrs() const override {
        return {&std::get<0>(b_), &std::get<1>(b_), &std::get<2>(b_),
                &std::get<3>(b_), &std::get<4>(b_), &std::get<5>(b_)};
    }
};

// TriangularPrism<Bac, Lef, Bot, Dia, Top> (subdomain.h:206-277)
template <typename Bac, typename Lef, typename Bot, typename Dia, typename Top = Bot>
class TriangularPrism : public Subdomain {
    std::tuple<Bac, Lef, Bot, Dia, Top> b_;

public:
    TriangularPrism(const Vector3d& o, const Matrix3d& mat, const Vector3l& div,
                    const Vector3d& gradT = Vector3d::Zero(),
                    const std::vector<double>& T = std::vector<double>(5, 0.0))
        : b_(Bac(o, Parallelogram(mat.col(1), mat.col(2)), T.at(0)),
             Lef(o, Parallelogram(mat.col(2), mat.col(0)), T.at(1)),
             Bot(o, Triangle(mat.col(0), mat.col(1)), T.at(2)),
             Dia(o + mat.col(0), Parallelogram(mat.col(2), mat.col(1) - mat.col(0)), T.at(3)),
             Top(o + mat.col(2), Triangle(mat.col(1), mat.col(0)), T.at(4))) {
        if (T.size() != 5) throw mcg::Error(MCG_EINVAL, "Incorrect number of temperatures");
        cell_ = MCG_TRIPRISM;
        o_ = o;
        mat_ = mat;
        div_ = div;
        gradT_ = gradT;
    }
    TriangularPrism(const TriangularPrism&) = delete;
    template <int I>
    typename std::tuple_element<I, std::tuple<Bac, Lef, Bot, Dia, Top>>::type& bdry() {
        return std::get<I>(b_);
    }
    std::vector<const Boundary*> bdryPtrs() const override {
        return {&std::get<0>(b_), &std::get<1>(b_), &std::get<2>(b_), &std::get<3>(b_),
                &std::get<4>(b_)};
    }
};

// Tetrahedron<Bac, Lef, Bot, Dia> (subdomain.h:285-352)
template <typename Bac, typename Lef, typename Bot, typename Dia>
class Tetrahedron : public Subdomain {
    std::tuple<Bac, Lef, Bot, Dia> b_;

public:
    Tetrahedron(const Vector3d& o, const Matrix3d& mat, const Vector3l& div,
                const Vector3d& gradT = Vector3d::Zero(),
                const std::vector<double>& T = std::vector<double>(4, 0.0))
        : b_(Bac(o, Triangle(mat.col(1), mat.col(2)), T.at(0)),
             Lef(o, Triangle(mat.col(2), mat.col(0)), T.at(1)),
             Bot(o, Triangle(mat.col(0), mat.col(1)), T.at(2)),
             Dia(o + mat.col(0), Triangle(mat.col(2) - mat.col(0), mat.col(1) - mat.col(0)), T.at(3))) {
        if (T.size() != 4) throw mcg::Error(MCG_EINVAL, "Incorrect number of temperatures");
        cell_ = MCG_TETRAHEDRON;
        o_ = o;
        mat_ = mat;
        div_ = div;
        gradT_ = gradT;
    }
    Tetrahedron(const Tetrahedron&) = delete;
    template <int I>
    typename std::tuple_element<I, std::tuple<Bac, Lef, Bot, Dia>>::type& bdry() {
        return std::get<I>(b_);
    }
    std::vector<const Boundary*> bdryPtrs() const override {
        return {&std::get<0>(b_), &std::get<1>(b_), &std::get<2>(b_), &std::get<3>(b_)};
    }
};

// Prism<Bot, Top, std::tuple<Sid0, ..., Sid(N-1)>>(o, mat, div, gradT, T)
// (subdomain.h:364-492; the reference takes a fusion::vector of sides): an
// N-gon base at o with vertices o + mat.col(1..N-1), the top at
// o + mat.col(0), side k a parallelogram from vertex k to vertex k + 1.
template <typename Bot, typename Top, typename Sid>
class Prism;
template <typename Bot, typename Top, typename... Sid>
class Prism<Bot, Top, std::tuple<Sid...>> : public Subdomain {
    static constexpr int N = int(sizeof...(Sid));
    static_assert(N >= 4 && N <= MCG_MAX_SIDES, "a prism base has 4 to 9 sides");
    using Cont = std::tuple<Bot, Top, Sid...>;
    Cont b_;

    static Matrix3Xd verts(const Matrix3Xd& mat, bool reversed) {
        Matrix3Xd v(3, N - 1);
        for (int k = 0; k < N - 1; ++k)
            for (int r = 0; r < 3; ++r) v(r, k) = mat(r, reversed ? N - 1 - k : k + 1);
        return v;
    }
    template <size_t... I>
    static Cont make(const Vector3d& o, const Matrix3Xd& mat, const std::vector<double>& T,
                     std::index_sequence<I...>) {
        return Cont(Bot(o, Polygon<N>(verts(mat, false)), T.at(0)),
                    Top(o + mat.col(0), Polygon<N>(verts(mat, true)), T.at(1)),
                    Sid(o + (I == 0 ? Vector3d::Zero() : mat.col(int(I))),
                        Parallelogram(mat.col(0), (I != N - 1 ? mat.col(int(I) + 1) : Vector3d::Zero()) -
                                                      (I == 0 ? Vector3d::Zero() : mat.col(int(I)))),
                        T.at(2 + I))...);
    }

public:
    Prism(const Vector3d& o, const Matrix3Xd& mat, long div, const Vector3d& gradT = Vector3d::Zero(),
          const std::vector<double>& T = std::vector<double>(N + 2, 0.0))
        : b_(make(o, check(mat, T), T, std::index_sequence_for<Sid...>{})) {
        cell_ = MCG_PRISM;
        o_ = o;
        matx_ = mat;
        div_ = Vector3l(div, 0, 0);
        gradT_ = gradT;
    }
    Prism(const Prism&) = delete;
    template <int I>
    typename std::tuple_element<I, Cont>::type& bdry() {
        return std::get<I>(b_);
    }
    std::vector<const Boundary*> bdryPtrs() const override {
        return ptrs(std::make_index_sequence<N + 2>{});
    }

private:
    static const Matrix3Xd& check(const Matrix3Xd& mat, const std::vector<double>& T) {
        if (mat.cols() != N) throw mcg::Error(MCG_EINVAL, "Incorrect number of matrix columns");
        if (T.size() != size_t(N + 2)) throw mcg::Error(MCG_EINVAL, "Incorrect number of temperatures");
        return mat;
    }
    template <size_t... I>
    std::vector<const Boundary*> ptrs(std::index_sequence<I...>) const {
        return {static_cast<const Boundary*>(&std::get<I>(b_))...};
    }
};

// Pyramid<Bot, std::tuple<Sid0, ..., Sid(N-1)>>(o, mat, div, gradT, T)
// (subdomain.h:504-630): an N-gon base at o, apex o + mat.col(0), side k a
// triangle from vertex k to vertex k + 1 and the apex.
template <typename Bot, typename Sid>
class Pyramid;
template <typename Bot, typename... Sid>
class Pyramid<Bot, std::tuple<Sid...>> : public Subdomain {
    static constexpr int N = int(sizeof...(Sid));
    static_assert(N >= 4 && N <= MCG_MAX_SIDES, "a pyramid base has 4 to 9 sides");
    using Cont = std::tuple<Bot, Sid...>;
    Cont b_;

    static Matrix3Xd verts(const Matrix3Xd& mat) {
        Matrix3Xd v(3, N - 1);
        for (int k = 0; k < N - 1; ++k)
            for (int r = 0; r < 3; ++r) v(r, k) = mat(r, k + 1);
        return v;
    }
    template <size_t... I>
    static Cont make(const Vector3d& o, const Matrix3Xd& mat, const std::vector<double>& T,
                     std::index_sequence<I...>) {
        return Cont(Bot(o, Polygon<N>(verts(mat)), T.at(0)),
                    Sid(o + (I == 0 ? Vector3d::Zero() : mat.col(int(I))),
                        Triangle(mat.col(0) - (I == 0 ? Vector3d::Zero() : mat.col(int(I))),
                                 (I != N - 1 ? mat.col(int(I) + 1) : Vector3d::Zero()) -
                                     (I == 0 ? Vector3d::Zero() : mat.col(int(I)))),
                        T.at(1 + I))...);
    }

public:
    Pyramid(const Vector3d& o, const Matrix3Xd& mat, long div, const Vector3d& gradT = Vector3d::Zero(),
            const std::vector<double>& T = std::vector<double>(N + 1, 0.0))
        : b_(make(o, check(mat, T), T, std::index_sequence_for<Sid...>{})) {
        cell_ = MCG_PYRAMID;
        o_ = o;
        matx_ = mat;
        div_ = Vector3l(div, 0, 0);
        gradT_ = gradT;
    }
    Pyramid(const Pyramid&) = delete;
    template <int I>
    typename std::tuple_element<I, Cont>::type& bdry() {
        return std::get<I>(b_);
    }
    std::vector<const Boundary*> bdryPtrs() const override {
        return ptrs(std::make_index_sequence<N + 1>{});
    }

private:
    static const Matrix3Xd& check(const Matrix3Xd& mat, const std::vector<double>& T) {
        if (mat.cols() != N) throw mcg::Error(MCG_EINVAL, "Incorrect number of matrix columns");
        if (T.size() != size_t(N + 1)) throw mcg::Error(MCG_EINVAL, "Incorrect number of temperatures");
        return mat;
    }
    template <size_t... I>
    std::vector<const Boundary*> ptrs(std::index_sequence<I...>) const {
        return {static_cast<const Boundary*>(&std::get<I>(b_))...};
    }
};

// ----------------------------------------------------------------- Domain
// Domain (domain.h:37-79): owns nothing here; derived domains own their cells
// and register them with addSdom in the reference's order.
class Domain {
    Subdomain::Pointers sdomPtrs_;
    std::vector<mcg_pore_spec> pores_;
    // the engine-layout description spec() assembles from the cells: a cache
    // behind a const accessor, so a Domain is used through const pointers as
    // the reference's problems hold it (problem.h:68)
    mutable std::vector<mcg_subdomain_spec> cells_;
    mutable std::vector<mcg_pair_spec> pairs_;
    mutable mcg_domain_spec spec_{};

    int locate(const Boundary* b, int* face) const {
        for (size_t s = 0; s < sdomPtrs_.size(); ++s) {
            std::vector<const Boundary*> bp = sdomPtrs_[s]->bdryPtrs();
            for (size_t f = 0; f < bp.size(); ++f)
                if (bp[f] == b) {
                    *face = int(f);
                    return int(s);
                }
        }
        throw mcg::Error(MCG_EINVAL, "paired face belongs to no subdomain of the domain");
    }

protected:
    void addSdom(const Subdomain* s) { sdomPtrs_.push_back(s); }
    void addPore(int sdom, double cx, double cy, double r, int kind = MCG_DIFF, double p = 0.0) {
        pores_.push_back(mcg_pore_spec{sdom, kind, {cx, cy}, r, p});
    }

public:
    virtual ~Domain() = default;
    Domain() = default;
    Domain(const Domain&) = delete;
    virtual std::string info() const = 0;
    const Subdomain::Pointers& sdomPtrs() const { return sdomPtrs_; }

    // The scene in the engine's layout: cells in registration order, one pair
    // spec per makePair call.
    const mcg_domain_spec& spec() const {
        cells_.clear();
        pairs_.clear();
        std::vector<std::pair<long, mcg_pair_spec>> ordered;
        for (const Subdomain* s : sdomPtrs_) cells_.push_back(s->spec());
        for (size_t s = 0; s < sdomPtrs_.size(); ++s) {
            std::vector<const Boundary*> bp = sdomPtrs_[s]->bdryPtrs();
            for (size_t f = 0; f < bp.size(); ++f) {
                std::vector<std::pair<const Boundary*, const Matrix3d*>> partners;
                std::vector<long> seqs;
                const Vector3d* t = nullptr;
                if (bp[f]->kind() == MCG_PERI) {
                    auto* pb = static_cast<const PeriBase*>(bp[f]);
                    if (pb->pair) {
                        partners.push_back({pb->pair, &pb->rot});
                        seqs.push_back(pb->pairSeq);
                        t = &pb->transl;
                    }
                } else if (bp[f]->kind() == MCG_INTER) {
                    auto* ib = static_cast<const InterBoundary*>(bp[f]);
                    for (size_t q = 0; q < ib->pairs.size(); ++q) {
                        partners.push_back({ib->pairs[q], nullptr});
                        seqs.push_back(ib->pairSeq[q]);
                    }
                }
                for (size_t q = 0; q < partners.size(); ++q) {
                    auto& pr = partners[q];
                    mcg_pair_spec ps{};
                    ps.sdom_a = int32_t(s);
                    ps.face_a = int32_t(f);
                    int fb = 0;
                    ps.sdom_b = locate(pr.first, &fb);
                    ps.face_b = fb;
                    if (t)
                        for (int k = 0; k < 3; ++k) ps.transl[k] = (*t)(k);
                    if (pr.second)
                        for (int i = 0; i < 9; ++i) ps.rot[i] = pr.second->m[i];
                    ordered.push_back({seqs[q], ps});
                }
            }
        }
        std::stable_sort(ordered.begin(), ordered.end(),
                         [](const auto& a, const auto& b) { return a.first < b.first; });
        for (auto& o : ordered) pairs_.push_back(o.second);
        spec_.nsdom = int32_t(cells_.size());
        spec_.npair = int32_t(pairs_.size());
        spec_.npore = int32_t(pores_.size());
        spec_.sdom = cells_.data();
        spec_.pair = pairs_.data();
        spec_.pore = pores_.data();
        return spec_;
    }
    virtual ArrayXXd average(const ArrayXXd& data) const { return data; }
    // points inside the domain that the check mode probes from (the
    // reference's checkpoints(), domain.h:68)
    virtual Matrix3Xd checkpoints() const { throw mcg::Error(MCG_EINVAL, "domain has no check points"); }
    friend std::ostream& operator<<(std::ostream& os, const Domain& d) { return os << d.info(); }
};

using PeriP = PeriBoundary<Parallelogram>;
using Spec = SpecBoundary;
using Diff = DiffBoundary;
using Inter = InterBoundary;
using IsotP = IsotBoundary<Parallelogram>;

// a 3 x n point list from n rows of (x, y, z)
inline Matrix3Xd points3(std::initializer_list<Vector3d> rows) {
    Matrix3Xd m(3, long(rows.size()));
    int k = 0;
    for (const Vector3d& v : rows) {
        for (int r = 0; r < 3; ++r) m(r, k) = v(r);
        ++k;
    }
    return m;
}

// BulkDomain (domain.h:93-114, domain.cpp:137-148)
class BulkDomain : public Domain {
public:
    using Sdom = Parallelepiped<PeriP, Spec, Spec>;

private:
    Vector3d dim_;
    Vector3l div_;
    double dT_;
    Sdom sdom_;

public:
    BulkDomain(const Vector3d& dim, const Vector3l& div, double dT)
        : dim_(dim), div_(div), dT_(dT),
          sdom_(Vector3d::Zero(), Matrix3d::Diagonal(dim), div, Vector3d(-dT / dim(0), 0., 0.)) {
        makePair(sdom_.bdry<0>(), sdom_.bdry<3>(), Vector3d(dim(0), 0., 0.));
        addSdom(&sdom_);
    }
    Matrix3Xd checkpoints() const override { return points3({dim_ * 0.5}); }   // domain.cpp:168-171
    std::string info() const override {
        std::ostringstream ss;
        ss << "BulkDomain " << this << "\n  dim: [" << dim_(0) << " " << dim_(1) << " " << dim_(2)
           << "]\n  div: [" << div_(0) << " " << div_(1) << " " << div_(2) << "]\n  dT:  " << dT_;
        return ss.str();
    }
};

// FilmDomain (domain.h:116-137, domain.cpp:188-199)
class FilmDomain : public Domain {
public:
    using Sdom = Parallelepiped<PeriP, Diff, Spec>;

private:
    Vector3d dim_;
    Vector3l div_;
    double dT_;
    Sdom sdom_;

public:
    FilmDomain(const Vector3d& dim, const Vector3l& div, double dT)
        : dim_(dim), div_(div), dT_(dT),
          sdom_(Vector3d::Zero(), Matrix3d::Diagonal(dim), div, Vector3d(-dT / dim(0), 0., 0.)) {
        makePair(sdom_.bdry<0>(), sdom_.bdry<3>(), Vector3d(dim(0), 0., 0.));
        addSdom(&sdom_);
    }
    Matrix3Xd checkpoints() const override { return points3({dim_ * 0.5}); }   // domain.cpp:219-222
    std::string info() const override {
        std::ostringstream ss;
        ss << "FilmDomain " << this << "\n  dim: [" << dim_(0) << " " << dim_(1) << " " << dim_(2)
           << "]\n  div: [" << div_(0) << " " << div_(1) << " " << div_(2) << "]\n  dT:  " << dT_;
        return ss.str();
    }
};

inline Matrix3d diag3(double a, double b, double c) { return Matrix3d::Diagonal(Vector3d(a, b, c)); }

// JctDomain (domain.h:184-220, domain.cpp:338-383)
class JctDomain : public Domain {
    Vector3d g_;
    Parallelepiped<PeriP, Spec, Diff, PeriP, Inter, Diff> s0_;
    Parallelepiped<PeriP, Inter, Diff, Inter, Spec, Diff> s1_;
    Parallelepiped<Inter, Inter, Diff, PeriP, Spec, Diff> s2_;
    std::array<double, 4> dim_;
    std::array<long, 4> div_;
    double dT_;

public:
    JctDomain(const std::array<double, 4>& dim, const std::array<long, 4>& div, double dT)
        : g_(-dT / (2. * dim[0]), 0., 0.),
          s0_(Vector3d(0., 0., 0.), diag3(2. * dim[0], dim[1], dim[3]), Vector3l(2 * div[0], div[1], div[3]), g_),
          s1_(Vector3d(0., dim[1], 0.), diag3(dim[0], dim[2], dim[3]), Vector3l(div[0], div[2], div[3]), g_),
          s2_(Vector3d(dim[0], dim[1], 0.), diag3(dim[0], dim[2], dim[3]), Vector3l(div[0], div[2], div[3]), g_),
          dim_(dim), div_(div), dT_(dT) {
        makePair(s0_.bdry<4>(), s1_.bdry<1>());
        makePair(s0_.bdry<4>(), s2_.bdry<1>());
        makePair(s1_.bdry<3>(), s2_.bdry<0>());
        const Vector3d transl(2. * dim[0], 0., 0.);
        makePair(s0_.bdry<0>(), s0_.bdry<3>(), transl);
        makePair(s1_.bdry<0>(), s2_.bdry<3>(), transl);
        addSdom(&s0_);
        addSdom(&s1_);
        addSdom(&s2_);
    }
    Matrix3Xd checkpoints() const override {   // domain.cpp:404-413
        const auto& d = dim_;
        return points3({Vector3d(d[0], 0.5 * d[1], 0.5 * d[3]), Vector3d(0.5 * d[0], d[1] + 0.5 * d[2], 0.5 * d[3]),
                        Vector3d(1.5 * d[0], d[1] + 0.5 * d[2], 0.5 * d[3])});
    }
    std::string info() const override {
        std::ostringstream ss;
        ss << "JctDomain " << this << "\n  dim: [" << dim_[0] << " " << dim_[1] << " " << dim_[2] << " "
           << dim_[3] << "]\n  div: [" << div_[0] << " " << div_[1] << " " << div_[2] << " " << div_[3]
           << "]\n  dT:  " << dT_;
        return ss.str();
    }
};

// TeeDomain (domain.h:222-259, domain.cpp:415-462)
class TeeDomain : public Domain {
    Vector3d g_;
    Parallelepiped<PeriP, Diff, Spec, Inter, Diff, Spec> s0_;
    Parallelepiped<Inter, Spec, Spec, Inter, Inter, Spec> s1_;
    Parallelepiped<Diff, Inter, Spec, Diff, Spec, Spec> s2_;
    Parallelepiped<Inter, Diff, Spec, PeriP, Diff, Spec> s3_;
    std::array<double, 5> dim_;
    std::array<long, 5> div_;
    double dT_;

public:
    TeeDomain(const std::array<double, 5>& dim, const std::array<long, 5>& div, double dT)
        : g_(-dT / (2. * dim[0] + dim[1]), 0., 0.),
          s0_(Vector3d(0., 0., 0.), diag3(dim[0], dim[2], dim[4]), Vector3l(div[0], div[2], div[4]), g_),
          s1_(Vector3d(dim[0], 0., 0.), diag3(dim[1], dim[2], dim[4]), Vector3l(div[1], div[2], div[4]), g_),
          s2_(Vector3d(dim[0], dim[2], 0.), diag3(dim[1], dim[3], dim[4]), Vector3l(div[1], div[3], div[4]), g_),
          s3_(Vector3d(dim[0] + dim[1], 0., 0.), diag3(dim[0], dim[2], dim[4]), Vector3l(div[0], div[2], div[4]),
              g_),
          dim_(dim), div_(div), dT_(dT) {
        makePair(s0_.bdry<3>(), s1_.bdry<0>());
        makePair(s1_.bdry<4>(), s2_.bdry<1>());
        makePair(s1_.bdry<3>(), s3_.bdry<0>());
        makePair(s0_.bdry<0>(), s3_.bdry<3>(), Vector3d(2. * dim[0] + dim[1], 0., 0.));
        addSdom(&s0_);
        addSdom(&s1_);
        addSdom(&s2_);
        addSdom(&s3_);
    }
    Matrix3Xd checkpoints() const override {   // domain.cpp:483-493
        const auto& d = dim_;
        return points3({Vector3d(0.5 * d[0], 0.5 * d[2], 0.5 * d[4]), Vector3d(d[0] + 0.5 * d[1], 0.5 * d[2], 0.5 * d[4]),
                        Vector3d(d[0] + 0.5 * d[1], d[2] + 0.5 * d[3], 0.5 * d[4]),
                        Vector3d(1.5 * d[0] + d[1], 0.5 * d[2], 0.5 * d[4])});
    }
    std::string info() const override {
        std::ostringstream ss;
        ss << "TeeDomain " << this << "\n  dim: [" << dim_[0] << " " << dim_[1] << " " << dim_[2] << " "
           << dim_[3] << " " << dim_[4] << "]\n  div: [" << div_[0] << " " << div_[1] << " " << div_[2] << " "
           << div_[3] << " " << div_[4] << "]\n  dT:  " << dT_;
        return ss.str();
    }
};

// TubeDomain (domain.h:261-297, domain.cpp:495-540)
class TubeDomain : public Domain {
    Vector3d g_;
    Parallelepiped<PeriP, Diff, Spec, PeriP, Diff, Inter> s0_;
    Parallelepiped<PeriP, Inter, Inter, PeriP, Diff, Diff> s1_;
    Parallelepiped<PeriP, Spec, Diff, PeriP, Inter, Diff> s2_;
    std::array<double, 4> dim_;
    std::array<long, 4> div_;
    double dT_;

public:
    TubeDomain(const std::array<double, 4>& dim, const std::array<long, 4>& div, double dT)
        : g_(-dT / dim[0], 0., 0.),
          s0_(Vector3d(0., dim[1], 0.), diag3(dim[0], dim[3], dim[2]), Vector3l(div[0], div[3], div[2]), g_),
          s1_(Vector3d(0., dim[1], dim[2]), diag3(dim[0], dim[3], dim[3]), Vector3l(div[0], div[3], div[3]), g_),
          s2_(Vector3d(0., 0., dim[2]), diag3(dim[0], dim[1], dim[3]), Vector3l(div[0], div[1], div[3]), g_),
          dim_(dim), div_(div), dT_(dT) {
        makePair(s1_.bdry<1>(), s2_.bdry<4>());
        makePair(s1_.bdry<2>(), s0_.bdry<5>());
        const Vector3d transl(dim[0], 0., 0.);
        makePair(s0_.bdry<0>(), s0_.bdry<3>(), transl);
        makePair(s1_.bdry<0>(), s1_.bdry<3>(), transl);
        makePair(s2_.bdry<0>(), s2_.bdry<3>(), transl);
        addSdom(&s0_);
        addSdom(&s1_);
        addSdom(&s2_);
    }
    Matrix3Xd checkpoints() const override {   // domain.cpp:561-570
        const auto& d = dim_;
        return points3({Vector3d(0.5 * d[0], d[1] + 0.5 * d[3], 0.5 * d[2]),
                        Vector3d(0.5 * d[0], d[1] + 0.5 * d[3], d[2] + 0.5 * d[3]),
                        Vector3d(0.5 * d[0], 0.5 * d[1], d[2] + 0.5 * d[3])});
    }
    std::string info() const override {
        std::ostringstream ss;
        ss << "TubeDomain " << this << "\n  dim: [" << dim_[0] << " " << dim_[1] << " " << dim_[2] << " "
           << dim_[3] << "]\n  div: [" << div_[0] << " " << div_[1] << " " << div_[2] << " " << div_[3]
           << "]\n  dT:  " << dT_;
        return ss.str();
    }
};

using Peri4 = PeriBoundary<Polygon<4>>;
template <typename... S>
using Sides = std::tuple<S...>;

inline Matrix3d cols3(const Vector3d& a, const Vector3d& b, const Vector3d& c) {
    Matrix3d m;
    const Vector3d* v[3] = {&a, &b, &c};
    for (int k = 0; k < 3; ++k)
        for (int r = 0; r < 3; ++r) m(r, k) = (*v[k])(r);
    return m;
}
inline Matrix3Xd cols4(const Vector3d& a, const Vector3d& b, const Vector3d& c, const Vector3d& d) {
    Matrix3Xd m(3, 4);
    const Vector3d* v[4] = {&a, &b, &c, &d};
    for (int k = 0; k < 4; ++k)
        for (int r = 0; r < 3; ++r) m(r, k) = (*v[k])(r);
    return m;
}

// OctetDomain (domain.h:299-350, domain.cpp:591-1225): one periodic unit of
// an octet nanotruss.  Nodes of 4-gon prisms, pyramids, triangular prisms and
// boxes, and ten struts (cells 16-25, the only sources and tallies), joined
// by 71 interface pairs and closed by 5 periodic pairs mapping x to -x.  The
// cells' geometry is the reference's (a table that has to be alike).
class OctetDomain : public Domain {
public:
    using SdomCont = std::tuple<
        Prism<Inter, Peri4, Sides<Inter, Diff, Spec, Diff>>,      // 00
        Prism<Diff, Inter, Sides<Inter, Diff, Spec, Inter>>,      // 01
        Prism<Diff, Diff, Sides<Inter, Inter, Spec, Spec>>,       // 02
        Prism<Inter, Peri4, Sides<Diff, Inter, Inter, Diff>>,     // 03
        Prism<Inter, Inter, Sides<Diff, Inter, Inter, Inter>>,    // 04
        Parallelepiped<Inter, Spec, Diff, Diff, Inter, Inter>,    // 05
        Prism<Spec, Inter, Sides<PeriP, Diff, Inter, Diff>>,      // 06
        Prism<Inter, Diff, Sides<PeriP, Diff, Inter, Inter>>,     // 07
        Prism<Diff, Diff, Sides<PeriP, Inter, Inter, Inter>>,     // 08
        Prism<Diff, Diff, Sides<Inter, Inter, Inter, Inter>>,     // 09
        Prism<Inter, Diff, Sides<Inter, Inter, Diff, Inter>>,     // 10
        Prism<Spec, Inter, Sides<Diff, Inter, Diff, Inter>>,      // 11
        TriangularPrism<Diff, Inter, Inter, Spec, Inter>,         // 12
        Pyramid<Inter, Sides<Diff, Inter, Spec, Inter>>,          // 13
        TriangularPrism<Inter, Inter, Diff, Spec, Inter>,         // 14
        Prism<Diff, Diff, Sides<Inter, Inter, Spec, Spec>>,       // 15
        Parallelepiped<Diff, Spec, Inter, Diff, Inter, Inter>,    // 16
        Parallelepiped<Inter, Inter, Inter, Diff, Diff, Inter>,   // 17
        Parallelepiped<Inter, Diff, Inter, Inter, Diff, Inter>,   // 18
        Parallelepiped<Diff, Inter, Inter, Inter, Diff, Inter>,   // 19
        Parallelepiped<Diff, Spec, Inter, Diff, Inter, Inter>,    // 20
        Parallelepiped<Diff, Spec, Inter, Diff, Inter, Inter>,    // 21
        Parallelepiped<Diff, Inter, Inter, Inter, Diff, Inter>,   // 22
        Parallelepiped<Inter, Diff, Inter, Inter, Diff, Inter>,   // 23
        Parallelepiped<Inter, Inter, Inter, Diff, Diff, Inter>,   // 24
        Parallelepiped<Diff, Spec, Inter, Diff, Inter, Inter>,    // 25
        Prism<Diff, Diff, Sides<Inter, Inter, Spec, Spec>>,       // 26
        TriangularPrism<Inter, Inter, Diff, Spec, Inter>,         // 27
        Pyramid<Inter, Sides<Diff, Inter, Spec, Inter>>,          // 28
        TriangularPrism<Diff, Inter, Inter, Spec, Inter>,         // 29
        Prism<Spec, Inter, Sides<Diff, Inter, Diff, Inter>>,      // 30
        Prism<Inter, Diff, Sides<Inter, Inter, Diff, Inter>>,     // 31
        Prism<Diff, Diff, Sides<Inter, Inter, Inter, Inter>>,     // 32
        Prism<Diff, Diff, Sides<PeriP, Inter, Inter, Inter>>,     // 33
        Prism<Inter, Diff, Sides<PeriP, Diff, Inter, Inter>>,     // 34
        Prism<Spec, Inter, Sides<PeriP, Diff, Inter, Diff>>,      // 35
        Parallelepiped<Inter, Spec, Diff, Diff, Inter, Inter>,    // 36
        Prism<Inter, Inter, Sides<Diff, Inter, Inter, Inter>>,    // 37
        Prism<Inter, Peri4, Sides<Diff, Inter, Inter, Diff>>,     // 38
        Prism<Diff, Diff, Sides<Inter, Inter, Spec, Spec>>,       // 39
        Prism<Diff, Inter, Sides<Inter, Diff, Spec, Inter>>,      // 40
        Prism<Inter, Peri4, Sides<Inter, Diff, Spec, Diff>>>;     // 41
    template <size_t I>
    using Cell = typename std::tuple_element<I, SdomCont>::type;

private:
    std::array<double, 4> dim_;
    std::array<long, 4> div_;
    double dT_;
    std::array<std::unique_ptr<Subdomain>, 42> cells_;

    template <size_t I, typename M, typename D>
    void make(const Vector3d& o, const M& mat, const D& div, const Vector3d& g = Vector3d::Zero()) {
        cells_[I].reset(new Cell<I>(o, mat, div, g));
    }

public:
    template <size_t I>
    Cell<I>& sdom() { return static_cast<Cell<I>&>(*cells_[I]); }

    OctetDomain(const std::array<double, 4>& dim, const std::array<long, 4>& div, double dT)
        : dim_(dim), div_(div), dT_(dT) {
        const double s = dim[0], t = dim[3], r2 = std::sqrt(2.0);
        const double pi = 3.141592653589793;          // constants.h:17
        const double a = pi / 5. * (dim[1] + dim[2]) - (4. - pi) / 5. * t, b = a / 4.;
        if (!(b > t) || !(a > (r2 + 1.) * (b + t)) || !(s / 2. > a + (r2 + 1.) * (b + t) + t))
            throw mcg::Error(MCG_EINVAL, "Invalid dimensions");
        const long n0 = div[0], n1 = div[1], n2 = div[2], n3 = div[3];
        const long no = -1;                           // noDiv(0)
        const Vector3l none(-1, -1, -1);              // noDiv
        const double gx = dT / (s - 2. * a - 2. * (r2 + 1.) * b - 2. * (2. + r2) * t);
        const Vector3d g(gx, 0., -gx);
        using V = Vector3d;
        // horizontal top: 00-05
        {
            const V c1(r2 * t, 0., 0.);
            const V c2(s / 4. - (2. + r2) / 2. * b - (2. - r2) / 2. * t, s / 4. - (2. + r2) / 2. * (b + t), 0.);
            const V c3(s / 4. - (2. + r2) / 2. * b - t, s / 4. - (2. + r2) / 2. * b - t, 0.);
            make<0>(V((r2 + 1.) * b + t, b + t, -a), cols4(V(0., 0., a), c1, c2, c3), no);
            make<1>(V((r2 + 1.) * b + t, b + t, -a - t), cols4(V(0., 0., t), c1, c2, c3), no);
            make<2>(V(b + t, b + t, -a - t),
                    cols4(V(0., 0., t), V(r2 * b, 0., 0.),
                          V(s / 4. - (2. - r2) / 2. * b - t, s / 4. - (2. + r2) / 2. * b - t, 0.),
                          V(s / 4. - b - t, s / 4. - b - t, 0.)), no);
            const V d1((r2 + 1.) * t, 0., 0.), d2((r2 + 1.) * t, t, 0.), d3(t, t, 0.);
            make<3>(V((r2 + 1.) * b, b, -a), cols4(V(0., 0., a), d1, d2, d3), no);
            make<4>(V((r2 + 1.) * b, b, -a - t), cols4(V(0., 0., t), d1, d2, d3), no);
            make<5>(V(b, b, -a - t), cols3(V(0., 0., t), V(r2 * b, 0., 0.), V(t, t, 0.)), none);
        }
        // node top: 06-15
        {
            const V c1(r2 * t, 0., 0.);
            const V c2((r2 + 1.) / 2. * b + (r2 + 1.) * t, 0., -(r2 + 1.) / 2. * b - t);
            const V c3((r2 + 1.) / 2. * b + (2. + r2) / 2. * t, 0., -(r2 + 1.) / 2. * b - (2. + r2) / 2. * t);
            make<6>(V(a, 0., 0.), cols4(V(0., b, 0.), c1, c2, c3), no);
            make<7>(V(a, b, 0.), cols4(V(0., t, 0.), c1, c2, c3), no);
            make<8>(V((r2 + 1.) * (b + t), b, 0.),
                    cols4(V(0., t, 0.), V(a - (r2 + 1.) * (b + t), 0., 0.),
                          V(a - (r2 + 1.) / 2. * b - t / r2, 0., -(r2 + 1.) / 2. * b - (2. + r2) / 2. * t),
                          V(0., 0., -a - t)), no);
            make<9>(V(b + t, b, -a - t),
                    cols4(V(0., t, 0.), V(r2 * (b + t), 0., 0.),
                          V((r2 - 1.) / 2. * b + t / r2, 0., -(r2 + 1.) / 2. * b - t / r2), V(0., 0., -b)), no);
            const V e1((r2 - 1.) / 2. * b + t / r2, 0., -(r2 - 1.) / 2. * b - t / r2);
            const V e2((r2 - 1.) / 2. * b, 0., -(r2 - 1.) / 2. * b - r2 * t), e3(0., 0., -r2 * t);
            make<10>(V(b + t, b, -a - b - t), cols4(V(0., t, 0.), e1, e2, e3), no);
            make<11>(V(b + t, 0., -a - b - t), cols4(V(0., b, 0.), e1, e2, e3), no);
            make<12>(V(b + t, b, -a - b), cols3(V(0., t, 0.), V(-t, 0., 0.), V(0., 0., b - t)), none);
            make<13>(V(b + t, b, -a - b - t), cols4(V(-t, 0., t), V(0., 0., t), V(0., t, t), V(0., t, 0.)), no);
            make<14>(V(b + t, b, -a - b - (r2 + 1.) * t), cols3(V(0., t, 0.), V(-t, 0., t), V(0., 0., r2 * t)),
                     none);
            make<15>(V(b + t, 0., -a - b - (r2 + 1.) * t),
                     cols4(V(0., 0., r2 * t), V(0., b, 0.), V(-t, b, t), V(-b - t, 0., b + t)), no);
        }
        // struts: 16-20 (top) and 21-25 (bottom)
        {
            const double W = (r2 + 1.) / 2. * b + (2. + r2) / 2. * t, K = s / 4. - a / 2. - W, h = t / r2;
            const V up(-K, 0., K), dn(K, 0., -K);
            make<16>(V(s / 4. + a / 2., 0., -s / 4. + a / 2.), cols3(V(h, 0., h), V(0., b, 0.), up),
                     Vector3l(n3, n2, n0), g);
            make<17>(V(s / 4. + a / 2., b, -s / 4. + a / 2.), cols3(V(h, 0., h), V(0., t, 0.), up),
                     Vector3l(n3, n3, n0), g);
            make<18>(V(s / 4. - a / 2., b, -s / 4. - a / 2.), cols3(V(a, 0., a), V(0., t, 0.), up),
                     Vector3l(n1, n3, n0), g);
            make<19>(V(s / 4. - a / 2. - h, b, -s / 4. - a / 2. - h), cols3(V(h, 0., h), V(0., t, 0.), up),
                     Vector3l(n3, n3, n0), g);
            make<20>(V(s / 4. - a / 2. - h, 0., -s / 4. - a / 2. - h), cols3(V(h, 0., h), V(0., b, 0.), up),
                     Vector3l(n3, n2, n0), g);
            make<21>(V(s / 4. + a / 2. + h, 0., -s / 4. + a / 2. + h), cols3(V(-h, 0., -h), V(0., b, 0.), dn),
                     Vector3l(n3, n2, n0), g);
            make<22>(V(s / 4. + a / 2. + h, b, -s / 4. + a / 2. + h), cols3(V(-h, 0., -h), V(0., t, 0.), dn),
                     Vector3l(n3, n3, n0), g);
            make<23>(V(s / 4. + a / 2., b, -s / 4. + a / 2.), cols3(V(-a, 0., -a), V(0., t, 0.), dn),
                     Vector3l(n1, n3, n0), g);
            make<24>(V(s / 4. - a / 2., b, -s / 4. - a / 2.), cols3(V(-h, 0., -h), V(0., t, 0.), dn),
                     Vector3l(n3, n3, n0), g);
            make<25>(V(s / 4. - a / 2., 0., -s / 4. - a / 2.), cols3(V(-h, 0., -h), V(0., b, 0.), dn),
                     Vector3l(n3, n2, n0), g);
        }
        // node bottom: 26-35
        {
            const double z0 = -s / 2. + a, x0 = s / 2. - b - t;
            make<26>(V(x0, 0., z0 + b + (r2 + 1.) * t),
                     cols4(V(0., 0., -r2 * t), V(0., b, 0.), V(t, b, -t), V(b + t, 0., -b - t)), no);
            make<27>(V(x0, b, z0 + b + (r2 + 1.) * t), cols3(V(0., t, 0.), V(t, 0., -t), V(0., 0., -r2 * t)), none);
            make<28>(V(x0, b, z0 + b + t), cols4(V(t, 0., -t), V(0., 0., -t), V(0., t, -t), V(0., t, 0.)), no);
            make<29>(V(x0, b, z0 + b), cols3(V(0., t, 0.), V(t, 0., 0.), V(0., 0., -b + t)), none);
            const V e1(-(r2 - 1.) / 2. * b - t / r2, 0., (r2 - 1.) / 2. * b + t / r2);
            const V e2(-(r2 - 1.) / 2. * b, 0., (r2 - 1.) / 2. * b + r2 * t), e3(0., 0., r2 * t);
            make<30>(V(x0, 0., z0 + b + t), cols4(V(0., b, 0.), e1, e2, e3), no);
            make<31>(V(x0, b, z0 + b + t), cols4(V(0., t, 0.), e1, e2, e3), no);
            make<32>(V(x0, b, z0 + t),
                     cols4(V(0., t, 0.), V(-r2 * (b + t), 0., 0.),
                           V(-(r2 - 1.) / 2. * b - t / r2, 0., (r2 + 1.) / 2. * b + t / r2), V(0., 0., b)), no);
            make<33>(V(s / 2. - (r2 + 1.) * (b + t), b, -s / 2.),
                     cols4(V(0., t, 0.), V(-a + (r2 + 1.) * (b + t), 0., 0.),
                           V(-a + (r2 + 1.) / 2. * b + t / r2, 0., (r2 + 1.) / 2. * b + (2. + r2) / 2. * t),
                           V(0., 0., a + t)), no);
            const V c1(-r2 * t, 0., 0.);
            const V c2(-(r2 + 1.) / 2. * b - (r2 + 1.) * t, 0., (r2 + 1.) / 2. * b + t);
            const V c3(-(r2 + 1.) / 2. * b - (2. + r2) / 2. * t, 0., (r2 + 1.) / 2. * b + (2. + r2) / 2. * t);
            make<34>(V(s / 2. - a, b, -s / 2.), cols4(V(0., t, 0.), c1, c2, c3), no);
            make<35>(V(s / 2. - a, 0., -s / 2.), cols4(V(0., b, 0.), c1, c2, c3), no);
        }
        // horizontal bottom: 36-41
        {
            const double z0 = -s / 2. + a, x1 = s / 2. - (r2 + 1.) * b;
            make<36>(V(s / 2. - b, b, z0 + t), cols3(V(0., 0., -t), V(-r2 * b, 0., 0.), V(-t, t, 0.)), none);
            const V d1(-(r2 + 1.) * t, 0., 0.), d2(-(r2 + 1.) * t, t, 0.), d3(-t, t, 0.);
            make<37>(V(x1, b, z0 + t), cols4(V(0., 0., -t), d1, d2, d3), no);
            make<38>(V(x1, b, z0), cols4(V(0., 0., -a), d1, d2, d3), no);
            make<39>(V(s / 2. - b - t, b + t, z0 + t),
                     cols4(V(0., 0., -t), V(-r2 * b, 0., 0.),
                           V(-s / 4. + (2. - r2) / 2. * b + t, s / 4. - (2. + r2) / 2. * b - t, 0.),
                           V(-s / 4. + b + t, s / 4. - b - t, 0.)), no);
            const V c1(-r2 * t, 0., 0.);
            const V c2(-s / 4. + (2. + r2) / 2. * b + (2. - r2) / 2. * t, s / 4. - (2. + r2) / 2. * (b + t), 0.);
            const V c3(-s / 4. + (2. + r2) / 2. * b + t, s / 4. - (2. + r2) / 2. * b - t, 0.);
            make<40>(V(x1 - t, b + t, z0 + t), cols4(V(0., 0., -t), c1, c2, c3), no);
            make<41>(V(x1 - t, b + t, z0), cols4(V(0., 0., -a), c1, c2, c3), no);
        }
        init();
        weigh();
        points(a, b, t);
    }
    OctetDomain(const OctetDomain&) = delete;
    Matrix3Xd checkpoints() const override { return pts_; }

    // OctetDomain::average (domain.cpp:1252-1280): one column, the mean of the
    // tallies over the struts' (cells 16-25) first layer along their third
    // axis, weighted by cell volume.
    ArrayXXd average(const ArrayXXd& data) const override {
        ArrayXXd out(data.rows(), 1);
        double wsum = 0.0;
        for (double w : weight_) wsum += w;
        for (long r = 0; r < data.rows(); ++r) {
            double acc = 0.0;
            for (long c = 0; c < data.cols(); ++c) acc += data(r, c) * weight_[size_t(c)];
            out(r, 0) = acc / wsum;
        }
        return out;
    }
    std::string info() const override {
        std::ostringstream ss;
        ss << "OctetDomain " << this << "\n  dim: [" << dim_[0] << " " << dim_[1] << " " << dim_[2] << " "
           << dim_[3] << "]\n  div: [" << div_[0] << " " << div_[1] << " " << div_[2] << " " << div_[3]
           << "]\n  dT:  " << dT_;
        return ss.str();
    }

private:
    // interface and periodic pairs (domain.cpp:595-687), then the cells in order
    void init() {
#define MCG_IP(A, FA, B, FB) makePair(sdom<A>().template bdry<FA>(), sdom<B>().template bdry<FB>())
        MCG_IP(0, 0, 1, 1); MCG_IP(1, 5, 2, 3); MCG_IP(0, 2, 3, 4); MCG_IP(1, 2, 4, 4);
        MCG_IP(2, 2, 5, 5); MCG_IP(3, 0, 4, 1); MCG_IP(4, 5, 5, 4);
        MCG_IP(3, 3, 8, 5); MCG_IP(4, 3, 8, 5); MCG_IP(4, 0, 9, 2); MCG_IP(5, 0, 9, 2); MCG_IP(5, 0, 12, 4);
        MCG_IP(6, 1, 7, 0); MCG_IP(7, 5, 8, 3); MCG_IP(9, 4, 10, 2); MCG_IP(10, 0, 11, 1);
        MCG_IP(9, 5, 12, 1); MCG_IP(9, 5, 13, 0); MCG_IP(10, 5, 14, 1); MCG_IP(11, 5, 15, 2);
        MCG_IP(12, 2, 13, 2); MCG_IP(13, 4, 14, 4); MCG_IP(14, 0, 15, 3);
        MCG_IP(6, 4, 16, 5); MCG_IP(7, 4, 17, 5); MCG_IP(8, 4, 18, 5); MCG_IP(9, 3, 18, 5);
        MCG_IP(10, 3, 19, 5); MCG_IP(11, 3, 20, 5);
        MCG_IP(16, 4, 17, 1); MCG_IP(17, 0, 18, 3); MCG_IP(18, 0, 19, 3); MCG_IP(19, 1, 20, 4);
        MCG_IP(16, 2, 21, 2); MCG_IP(17, 2, 22, 2); MCG_IP(18, 2, 23, 2); MCG_IP(19, 2, 24, 2);
        MCG_IP(20, 2, 25, 2);
        MCG_IP(21, 4, 22, 1); MCG_IP(22, 3, 23, 0); MCG_IP(23, 3, 24, 0); MCG_IP(24, 1, 25, 4);
        MCG_IP(21, 5, 30, 3); MCG_IP(22, 5, 31, 3); MCG_IP(23, 5, 32, 3); MCG_IP(23, 5, 33, 4);
        MCG_IP(24, 5, 34, 4); MCG_IP(25, 5, 35, 4);
        MCG_IP(26, 3, 27, 0); MCG_IP(27, 4, 28, 4); MCG_IP(28, 2, 29, 2); MCG_IP(26, 2, 30, 5);
        MCG_IP(27, 1, 31, 5); MCG_IP(28, 0, 32, 5); MCG_IP(29, 1, 32, 5); MCG_IP(30, 1, 31, 0);
        MCG_IP(31, 2, 32, 4); MCG_IP(33, 3, 34, 5); MCG_IP(34, 0, 35, 1);
        MCG_IP(29, 4, 36, 0); MCG_IP(32, 2, 36, 0); MCG_IP(32, 2, 37, 0); MCG_IP(33, 5, 37, 3);
        MCG_IP(33, 5, 38, 3);
        MCG_IP(36, 4, 37, 5); MCG_IP(37, 1, 38, 0); MCG_IP(36, 5, 39, 2); MCG_IP(37, 4, 40, 2);
        MCG_IP(38, 4, 41, 2); MCG_IP(39, 3, 40, 5); MCG_IP(40, 1, 41, 0);
#undef MCG_IP
        const Vector3d transl(dim_[0] / 2., 0., -dim_[0] / 2.);
        const Matrix3d rot = diag3(-1., 1., 1.);
        makePair(sdom<0>().bdry<1>(), sdom<41>().bdry<1>(), transl, rot);
        makePair(sdom<3>().bdry<1>(), sdom<38>().bdry<1>(), transl, rot);
        makePair(sdom<6>().bdry<2>(), sdom<35>().bdry<2>(), transl, rot);
        makePair(sdom<7>().bdry<2>(), sdom<34>().bdry<2>(), transl, rot);
        makePair(sdom<8>().bdry<2>(), sdom<33>().bdry<2>(), transl, rot);
        for (auto& c : cells_) addSdom(c.get());
    }

    // the 52 check points of domain.cpp:693-1221, from each cell's origin o
    // and edge columns c: mostly o + (c0 + c2) / 2, with the reference's
    // other rules for the cells named below
    Matrix3Xd pts_;
    void points(double a, double b, double t) {
        const double r2 = std::sqrt(2.0), h = t / r2;
        std::vector<Vector3d> p;
        for (int i = 0; i < 42; ++i) {
            const mcg_subdomain_spec c = cells_[size_t(i)]->spec();
            const Vector3d o(c.origin[0], c.origin[1], c.origin[2]);
            auto col = [&](int k) { return Vector3d(c.mat[3 * k], c.mat[3 * k + 1], c.mat[3 * k + 2]); };
            const Vector3d c0 = col(0), c1 = col(1), c2 = col(2);
            switch (i) {
                case 5: case 36: {
                    const Vector3d u(i == 5 ? t : -t, 0., 0.);
                    p.push_back(o + c0 * 0.5 + (c2 + u) * (1. / 3.));
                    p.push_back(o + (c0 + c1 + c2 + u) * 0.5);
                    break;
                }
                case 8: case 33: {
                    const Vector3d v(0., 0., i == 8 ? -a : a);
                    p.push_back(o + (c0 + c1 + v) * 0.5);
                    p.push_back(o + (c0 + c2 + v) * 0.5);
                    break;
                }
                case 9: case 32: {
                    const Vector3d u(i == 9 ? r2 * b : -r2 * b, 0., 0.), v(0., 0., i == 9 ? -b + t : b - t);
                    p.push_back(o + (c0 + c2 + u) * 0.5);
                    p.push_back(o + (c0 + c2 + v) * 0.5);
                    p.push_back(o + (c0 + u + v) * 0.5);
                    break;
                }
                case 12: case 14: case 27: case 29:
                    p.push_back(o + (c0 + c1) * (1. / 3.) + c2 * 0.5);
                    break;
                case 13: case 28:
                    p.push_back(o + (c0 + c1 + c2 + col(3)) * 0.2);
                    break;
                case 18: case 23: {
                    const double e = (r2 + 1.) / 2. * b + h, sg = i == 18 ? 1. : -1.;
                    const Vector3d u(sg * e, 0., sg * e);
                    p.push_back(o + (c1 + c2 + u) * 0.5);
                    p.push_back(o + (c0 + c1 + c2 + u) * 0.5);
                    break;
                }
                default:
                    if (i >= 16 && i < 26) p.push_back(o + (c0 + c1 + c2) * 0.5);
                    else p.push_back(o + (c0 + c2) * 0.5);
            }
        }
        pts_ = Matrix3Xd(3, long(p.size()));
        for (size_t k = 0; k < p.size(); ++k)
            for (int r = 0; r < 3; ++r) pts_(r, int(k)) = p[k](r);
    }

    // WeightF (domain.cpp:1260-1280) over Field's columns: each cell's
    // columns run with the first index fastest (field.cpp:25-45)
    std::vector<double> weight_;
    void weigh() {
        const mcg_domain_spec& d = spec();
        int64_t n = 0;
        mcg::check(mcg_cell_volumes(&d, nullptr, 0, &n));
        std::vector<double> vol(size_t(n), 0.0);
        mcg::check(mcg_cell_volumes(&d, vol.data(), n, &n));
        weight_.assign(size_t(n), 0.0);
        int64_t col = 0;
        for (int i = 0; i < d.nsdom; ++i) {
            const mcg_subdomain_spec& c = d.sdom[i];
            int64_t sh[3] = {1, 1, 1};
            if (c.cell == MCG_PRISM || c.cell == MCG_PYRAMID) {
                if (c.div[0] < 0) sh[0] = sh[1] = sh[2] = 0;
            } else if (c.div[0] < 0 || c.div[1] < 0 || c.div[2] < 0) {
                sh[0] = sh[1] = sh[2] = 0;
            } else {
                for (int k = 0; k < 3; ++k) sh[k] = std::max<int64_t>(c.div[k], 1);
            }
            const int64_t per_layer = sh[0] * sh[1], total = per_layer * sh[2];
            if (i >= 16 && i < 26)
                for (int64_t q = 0; q < per_layer; ++q) weight_[size_t(col + q)] = vol[size_t(col + q)];
            col += total;
        }
    }
};

// Periodic porous film (an extension: the reference's boundaries are planes).
// One box cell periodic in x, with the volumetric gradient source -dT/dim0
// of FilmDomain (domain.cpp:188-199), and in y; walls at z = 0 and z = dim2;
// a cylindrical pore along z of radius r centred in the x-y cross-section.
template <typename Wall = Diff>
class PorousFilmDomain : public Domain {
public:
    using Sdom = Parallelepiped<PeriP, PeriP, Wall>;

private:
    Vector3d dim_;
    Vector3l div_;
    double radius_, dT_;
    Sdom sdom_;

public:
    PorousFilmDomain(const Vector3d& dim, double radius, const Vector3l& div, double dT,
                     int poreKind = MCG_DIFF, double specularity = 0.0)
        : dim_(dim), div_(div), radius_(radius), dT_(dT),
          sdom_(Vector3d::Zero(), Matrix3d::Diagonal(dim), div, Vector3d(-dT / dim(0), 0., 0.)) {
        makePair(sdom_.template bdry<0>(), sdom_.template bdry<3>(), Vector3d(dim(0), 0., 0.));
        makePair(sdom_.template bdry<1>(), sdom_.template bdry<4>(), Vector3d(0., dim(1), 0.));
        addSdom(&sdom_);
        addPore(0, 0.5 * dim(0), 0.5 * dim(1), radius, poreKind, specularity);
    }
    std::string info() const override {
        std::ostringstream ss;
        ss << "PorousFilmDomain " << this << "\n  dim: [" << dim_(0) << " " << dim_(1) << " " << dim_(2)
           << "]\n  pore radius: " << radius_ << "\n  div: [" << div_(0) << " " << div_(1) << " "
           << div_(2) << "]\n  dT:  " << dT_;
        return ss.str();
    }
};

// Cross-plane slab / bar with isothermal ends (the workloads' domains; the
// reference ships none with IsotBoundary, the composition is API-legal).
template <typename Side>
class WallDomain : public Domain {
public:
    using Sdom = Parallelepiped<IsotP, Side, Side>;

private:
    Vector3d dim_;
    Sdom sdom_;

public:
    WallDomain(const Vector3d& dim, const Vector3l& div, double Thot, double Tcold)
        : dim_(dim),
          sdom_(Vector3d::Zero(), Matrix3d::Diagonal(dim), div, Vector3d::Zero(),
                std::vector<double>{Thot, 0., 0., Tcold, 0., 0.}) {
        addSdom(&sdom_);
    }
    Sdom& sdom() { return sdom_; }
    std::string info() const override {
        std::ostringstream ss;
        ss << "WallDomain " << this << "\n  dim: [" << dim_(0) << " " << dim_(1) << " " << dim_(2) << "]";
        return ss.str();
    }
};
using SlabDomain = WallDomain<Spec>;
using BarDomain = WallDomain<Diff>;

// ------------------------------------------------------------------ Field
// Field(rows, dom) (field.h:26-49): the rows x cells tally layout, one block
// of columns per subdomain with a grid, in registration order (field.cpp:25-45).
class Field {
    ArrayXXd data_;

public:
    Field(long rows, const Domain* dom) {
        const mcg_domain_spec& d = dom->spec();
        long cols = 0;
        for (int s = 0; s < d.nsdom; ++s) {
            const int64_t* dv = d.sdom[s].div;
            if (dv[0] < 0 || dv[1] < 0 || dv[2] < 0) continue;
            cols += std::max<int64_t>(dv[0], 1) * std::max<int64_t>(dv[1], 1) * std::max<int64_t>(dv[2], 1);
        }
        data_ = ArrayXXd(rows, cols);
    }
    const ArrayXXd& data() const { return data_; }
};

// ------------------------------------------------------------- Progress
class Progress {
    long tot_ = 0, count_ = 0, esc_ = 0;

public:
    Progress() = default;
    Progress(long tot, long) : tot_(tot) {}
    long count() const { return count_; }
    long esc() const { return esc_; }
    long incrCount() { return ++count_; }
    long incrEsc() { return ++esc_; }
    void add(long bundles, long esc) {
        count_ += bundles;
        esc_ += esc;
    }
};

// ----------------------------------------------------------- Statistics
// Welford mean / sample variance (field.cpp:226-262)
template <typename S>
class Statistics {
    long n_ = 0;
    S z_, m_, s_;

public:
    Statistics() = default;
    explicit Statistics(const S& zero) : z_(zero), m_(zero), s_(zero) {}
    S mean() const { return m_; }
    S variance() const { return n_ < 2 ? z_ : s_ / double(n_ - 1); }
    void add(const S& x) {
        n_++;
        S d = x - m_;
        m_ += d / double(n_);
        s_ += d * (x - m_);
    }
};

// ---------------------------------------------------------------- Problem
class Problem {
    const Material* mat_;
    const Domain* dom_;

public:
    Problem(const Material* mat, const Domain* dom) : mat_(mat), dom_(dom) {}   // problem.h:76
    virtual ~Problem() = default;
    const Material* mat() const { return mat_; }
    const Domain* dom() const { return dom_; }
    virtual Progress initProgress() const = 0;
    virtual ArrayXXd solve(Rng& gen, Progress* prog) const = 0;
};

// TrajProblem(mat, dom, [prop,] pos, [dir,] maxscat, maxloop) with a given
// start (problem.h:88-119, problem.cpp:163-299): one trajectory on the GPU
// (mcg_trace_from).  solve() returns the points the bundle visits, 3 x n: the
// start, then every position it takes (the end of each flight, and the
// mapped position after a periodic face, as TrkPhonon records them); hops()
// holds the step records.
class TrajProblem : public Problem {
    mutable mcg_scene* scene_ = nullptr;
    long maxscat_, maxloop_;
    long prop_ = -1;                  // (w, p) cell w * np + p, or -1: drawn
    Vector3d pos_, dir_;
    bool hasDir_;
    int precision_ = MCG_FP64;
    mutable std::vector<double> rec_;
    mutable uint64_t lastKey_ = 0;

    static constexpr int64_t kMaxRecords = int64_t(1) << 20;   // steps one solve() records at most

    static long propCell(const Material* mat, const Phonon::Prop& prop) {
        if (prop.w() < 0 || prop.w() >= mat->nw() || prop.p() < 0 || prop.p() >= mat->np())
            throw mcg::Error(MCG_EINVAL, "prop outside the material's table");
        return prop.w() * mat->np() + prop.p();
    }

public:
    TrajProblem(const Material* mat, const Domain* dom, const Phonon::Prop& prop, const Vector3d& pos,
                const Vector3d& dir, long maxscat = 100, long maxloop = 0)
        : Problem(mat, dom), maxscat_(maxscat), maxloop_(maxloop), prop_(propCell(mat, prop)), pos_(pos),
          dir_(dir), hasDir_(true) {}
    TrajProblem(const Material* mat, const Domain* dom, const Phonon::Prop& prop, const Vector3d& pos,
                long maxscat = 100, long maxloop = 0)
        : Problem(mat, dom), maxscat_(maxscat), maxloop_(maxloop), prop_(propCell(mat, prop)), pos_(pos),
          hasDir_(false) {}
    TrajProblem(const Material* mat, const Domain* dom, const Vector3d& pos, const Vector3d& dir,
                long maxscat = 100, long maxloop = 0)
        : Problem(mat, dom), maxscat_(maxscat), maxloop_(maxloop), pos_(pos), dir_(dir), hasDir_(true) {}
    TrajProblem(const Material* mat, const Domain* dom, const Vector3d& pos, long maxscat = 100,
                long maxloop = 0)
        : Problem(mat, dom), maxscat_(maxscat), maxloop_(maxloop), pos_(pos), hasDir_(false) {}
    TrajProblem(const TrajProblem&) = delete;
    ~TrajProblem() override {
        if (scene_) mcg_scene_destroy(scene_);
    }
    void precision(int p) { precision_ = p; }
    uint64_t lastKey() const { return lastKey_; }
    Progress initProgress() const override { return Progress(); }
    ArrayXXd solve(Rng& gen, Progress*) const override {
        if (!scene_) {
            mcg_material_spec ms = mat()->spec();
            mcg::check(mcg_scene_create(&ms, &dom()->spec(), &scene_));
        }
        mcg_problem_spec ps{};
        ps.tally = MCG_TEMP;
        ps.precision = precision_;
        ps.nemit = 1;
        ps.maxscat = maxscat_;
        ps.maxloop = maxloop_;
        lastKey_ = gen();
        ps.seed = lastKey_;
        const int64_t loops = maxloop_ != 0 ? maxloop_ : 100 * maxscat_;
        const int64_t cap = std::max<int64_t>(1, std::min<int64_t>(loops, kMaxRecords));
        rec_.assign(size_t(cap) * MCG_TRACE_WIDTH, 0.0);
        int64_t n = 0;
        const double p[3] = {pos_(0), pos_(1), pos_(2)}, d[3] = {dir_(0), dir_(1), dir_(2)};
        mcg::check(mcg_trace_from(scene_, &ps, prop_, p, hasDir_ ? d : nullptr, rec_.data(), cap, &n));
        rec_.resize(size_t(n) * MCG_TRACE_WIDTH);
        if (n == cap && rec_[size_t(n - 1) * MCG_TRACE_WIDTH + 13] < 0.0)
            throw mcg::Error(MCG_ERANGE, "trajectory longer than the record buffer (" +
                                             std::to_string(kMaxRecords) + " steps)");
        // the faces' kinds in the engine's plane order (cells in order, each
        // cell's faces in container order)
        const mcg_domain_spec& ds = dom()->spec();
        std::vector<int> kind;
        for (int si = 0; si < ds.nsdom; ++si)
            for (int f = 0; f < ds.sdom[si].nface; ++f) kind.push_back(ds.sdom[si].face[f].kind);
        // TrkPhonon records a point whenever the position changes
        // (phonon.cpp:154-157): the end of every flight (Phonon::move), and
        // the mapped position after a periodic face (boundary.cpp:519), which
        // a flight that ended the bundle outside its cell never reaches
        std::vector<double> pts{pos_(0), pos_(1), pos_(2)};
        for (int64_t i = 0; i < n; ++i) {
            const double* r = rec_.data() + size_t(i) * MCG_TRACE_WIDTH;
            pts.insert(pts.end(), r + 3, r + 6);
            const long hit = long(r[1]);
            if (hit >= 0 && hit < long(kind.size()) && kind[size_t(hit)] == MCG_PERI && r[13] != double(MCG_ESCAPED))
                pts.insert(pts.end(), r + 14, r + 17);
        }
        const long np = long(pts.size() / 3);
        ArrayXXd traj(3, np);
        for (long c = 0; c < np; ++c)
            for (int k = 0; k < 3; ++k) traj(k, c) = pts[size_t(3 * c + k)];
        return traj;
    }
    // the step records of the last solve (MCG_TRACE_WIDTH doubles each)
    const std::vector<double>& hops() const { return rec_; }
    std::string info() const {
        std::ostringstream ss;
        ss << "TrajProblem " << this << "\n  pos:     [" << pos_(0) << " " << pos_(1) << " " << pos_(2) << "]";
        if (hasDir_) ss << "\n  dir:     [" << dir_(0) << " " << dir_(1) << " " << dir_(2) << "]";
        ss << "\n  maxscat: " << maxscat_ << "\n  maxloop: " << maxloop_;
        return ss.str();
    }
    friend std::ostream& operator<<(std::ostream& os, const TrajProblem& p) { return os << p.info(); }
};

// FieldProblem(mat, dom, nemit, maxscat, maxloop) (problem.cpp:315-342).
// The engine rounds the per-emitter counts and computes power_; solve() runs
// one GPU solve and returns power_ * postProc(data) / cellVol.
class FieldProblem : public Problem {
    mutable mcg_scene* scene_ = nullptr;
    int tally_;

protected:
    long nemit_, maxscat_, maxloop_, size_ = 0;
    int precision_ = MCG_FP32_FLIGHT;
    mutable mcg_problem_info info_{};
    mutable mcg_stats last_{};
    mutable uint64_t lastKey_ = 0;

    mcg_problem_spec problemSpec(uint64_t seed) const {
        mcg_problem_spec p{};
        p.tally = tally_;
        p.precision = precision_;
        p.nemit = nemit_;
        p.maxscat = maxscat_;
        p.maxloop = maxloop_;
        p.size = size_;
        p.seed = seed;
        return p;
    }

public:
    FieldProblem(const Material* mat, const Domain* dom, int tally, long nemit, long maxscat,
                 long maxloop)
        : Problem(mat, dom), tally_(tally), nemit_(nemit), maxscat_(maxscat), maxloop_(maxloop) {}
    ~FieldProblem() override {
        if (scene_) mcg_scene_destroy(scene_);
    }
    void precision(int p) { precision_ = p; }
    const mcg_problem_info& plan() const {
        mcg_material_spec ms = mat()->spec();
        mcg_problem_spec ps = problemSpec(0);
        mcg::check(mcg_problem_plan(&ms, &dom()->spec(), &ps, &info_, nullptr, nullptr, 0));
        return info_;
    }
    long nemit() const { return plan().nemit; }
    double power() const { return plan().power; }
    Progress initProgress() const override { return Progress(plan().nemit, std::min(20l, plan().nemit)); }
    ArrayXXd initSolution() const { return ArrayXXd(plan().rows, info_.cols); }
    const mcg_stats& lastStats() const { return last_; }
    uint64_t lastKey() const { return lastKey_; }     // Philox key of the last solve

    ArrayXXd solve(Rng& gen, Progress* prog) const override {
        const mcg_problem_info& info = plan();
        if (!scene_) {
            mcg_material_spec ms = mat()->spec();
            mcg::check(mcg_scene_create(&ms, &dom()->spec(), &scene_));
        }
        ArrayXXd out(info.rows, info.cols);
        lastKey_ = gen();
        mcg_problem_spec ps = problemSpec(lastKey_);
        mcg::check(mcg_solve(scene_, &ps, out.data(), out.size(), nullptr, &last_));
        if (prog) prog->add(last_.bundles, last_.escapes);
        return out;
    }
};

struct TempProblem : FieldProblem {
    TempProblem(const Material* m, const Domain* d, long nemit, long maxscat, long maxloop = 0)
        : FieldProblem(m, d, MCG_TEMP, nemit, maxscat, maxloop) {}
};
struct FluxProblem : FieldProblem {
    FluxProblem(const Material* m, const Domain* d, long nemit, long maxscat, long maxloop = 0)
        : FieldProblem(m, d, MCG_FLUX, nemit, maxscat, maxloop) {}
};
struct MultiProblem : FieldProblem {
    MultiProblem(const Material* m, const Domain* d, long nemit, long maxscat, long maxloop = 0)
        : FieldProblem(m, d, MCG_MULTI, nemit, maxscat, maxloop) {}
};
struct CumTempProblem : FieldProblem {
    CumTempProblem(const Material* m, const Domain* d, long nemit, long size, long maxscat, long maxloop = 0)
        : FieldProblem(m, d, MCG_CUMTEMP, nemit, maxscat, maxloop) {
        size_ = size;
    }
};
struct CumFluxProblem : FieldProblem {
    CumFluxProblem(const Material* m, const Domain* d, long nemit, long size, long maxscat, long maxloop = 0)
        : FieldProblem(m, d, MCG_CUMFLUX, nemit, maxscat, maxloop) {
        size_ = size;
    }
};

// ------------------------------------------------------------ driver glue
// printSolution (main.cpp:67-84): "%16.9e " per entry, one row per line.
inline void printSolution(const ArrayXXd& sol, std::ostream& os = std::cout) {
    char buf[64];
    for (long i = 0; i < sol.rows(); ++i) {
        for (long j = 0; j < sol.cols(); ++j) {
            std::snprintf(buf, sizeof buf, "%16.9e ", sol(i, j));
            os << buf;
        }
        os << "\n";
    }
    os << "\n";
}

// solveField (main.cpp:146-179): one GPU solve (the reference sums one
// partial per OpenMP thread; the engine already covers every bundle).  The
// log gets the reference's "Solution n" and "Output" blocks; the line after
// the header gives the solve's Philox key (where the reference prints its
// threads' seeds).
inline ArrayXXd solveField(const FieldProblem* prob, Rng& gen, std::ostream* log = nullptr) {
    static long n = 0;
    if (log) *log << "Solution " << n++ << "\n";
    Progress prog = prob->initProgress();
    ArrayXXd sol = prob->solve(gen, &prog);
    if (log) {
        *log << "  key: " << prob->lastKey() << "\n";
        *log << "Output\n";
        printSolution(sol, *log);
    }
    ArrayXXd avg = prob->dom()->average(sol);
    bool same = avg.rows() == sol.rows() && avg.cols() == sol.cols();
    for (long i = 0; same && i < sol.size(); ++i) same = avg.data()[i] == sol.data()[i];
    if (log && !same) {
        *log << "Averaged\n";
        printSolution(avg, *log);
    }
    return avg;
}

// solveField(prob, nsim) + calcStats (main.cpp:181-210)
inline Statistics<ArrayXXd> calcStats(const FieldProblem* prob, long nsim, Rng& gen,
                                      std::ostream* log = nullptr) {
    Statistics<ArrayXXd> stats(prob->initSolution());
    for (long i = 0; i < nsim; ++i) stats.add(solveField(prob, gen, log));
    if (log) {
        *log << "Mean\n";
        printSolution(stats.mean(), *log);
        *log << "Standard Deviation\n";
        printSolution(stats.variance().sqrt(), *log);
    }
    return stats;
}
