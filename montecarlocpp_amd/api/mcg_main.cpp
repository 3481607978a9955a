// mcg_main.cpp -- the reference's command line (main.cpp:216-479) on the GPU.
//
//   mcg_main <prefix> <material> <domain> <problem>
//
//   material:  grey [T] | silicon [T] | custom [T] [disp] [relax]
//              (files read from <prefix>, as main.cpp:257-281 does)
//   domain:    bulk [dim] [div0]           BulkDomain, dT = 1e6 dim
//              film [dim0,2] [dim1] [div1] FilmDomain, dT = 1e6 dim0
//              slab [L] [W] [div] [Thot] [Tcold]   isothermal walls, specular sides
//              bar  [L] [W] [div] [Thot] [Tcold]   isothermal ends, diffuse sides
//              jct [d0] [d3] | tee [d0] [d4] [div] | tube [d0] [d1] [d3] [div1] [div3]
//                         the reference's assemblies (main.cpp:339-375)
//              porous [side] [thick] [radius] [div]  periodic porous film: side x
//                         side x thick cell, centred pore, dT = 1 K across x,
//                         diffuse pore, top and bottom, div x div map
//   problem:   temp|flux|multi [nemit] [maxscat] [maxloop] [nsim]
//              cumtemp|cumflux [nemit] [size] [maxscat] [maxloop] [nsim]
//
// Environment: MCG_SEED (the Rng seed; default std::random_device),
// MCG_PRECISION=fp64|fp32 (default fp32 flight), MCG_DUMP_SCENE=1 prints the
// flattened scene as JSON and exits without touching the GPU.
// Output follows main.cpp: the argument echo, object info, "Solution n",
// "Output" and the solution matrix in %16.9e, then Mean / Standard Deviation
// when nsim > 1.
#include <unistd.h>

#include <chrono>
#include <cstdlib>
#include <ctime>
#include <iomanip>
#include <limits>

#include "montecarlocpp.hpp"

namespace {

void dump_scene(const mcg_domain_spec& d) {
    std::cout << std::setprecision(17) << "{\"cells\": [";
    for (int s = 0; s < d.nsdom; ++s) {
        const mcg_subdomain_spec& c = d.sdom[s];
        std::cout << (s ? ", " : "") << "{\"cell\": " << c.cell << ", \"origin\": [" << c.origin[0]
                  << ", " << c.origin[1] << ", " << c.origin[2] << "], \"mat\": [";
        for (int i = 0; i < 3 * MCG_MAX_SIDES; ++i) std::cout << (i ? ", " : "") << c.mat[i];
        std::cout << "], \"div\": [" << c.div[0] << ", " << c.div[1] << ", " << c.div[2]
                  << "], \"grad_t\": [" << c.grad_t[0] << ", " << c.grad_t[1] << ", " << c.grad_t[2]
                  << "], \"faces\": [";
        for (int f = 0; f < c.nface; ++f)
            std::cout << (f ? ", " : "") << "[" << c.face[f].kind << ", " << c.face[f].temp << ", "
                      << c.face[f].specularity << "]";
        std::cout << "]}";
    }
    std::cout << "], \"pairs\": [";
    for (int k = 0; k < d.npair; ++k) {
        const mcg_pair_spec& p = d.pair[k];
        std::cout << (k ? ", " : "") << "[" << p.sdom_a << ", " << p.face_a << ", " << p.sdom_b
                  << ", " << p.face_b << ", " << p.transl[0] << ", " << p.transl[1] << ", "
                  << p.transl[2] << "]";
    }
    std::cout << "], \"pores\": [";
    for (int k = 0; k < d.npore; ++k) {
        const mcg_pore_spec& q = d.pore[k];
        std::cout << (k ? ", " : "") << "[" << q.sdom << ", " << q.kind << ", " << q.center[0] << ", "
                  << q.center[1] << ", " << q.radius << ", " << q.specularity << "]";
    }
    std::cout << "]}" << std::endl;
}

std::string timestamp() {
    std::time_t now = std::time(nullptr);
    char buf[80];
    std::strftime(buf, sizeof buf, "%Y-%m-%d %H:%M:%S", std::localtime(&now));
    return buf;
}

}  // namespace

int main(int argc, const char* argv[]) {
    std::stringstream argss;
    for (int i = 1; i < argc; ++i) argss << argv[i] << ' ';
    std::cout << std::string(40, '-') << "\n" << timestamp() << "\n" << argss.str() << "\n\n";
    try {
        std::string prefix, matStr, domStr, probStr;
        argss >> prefix >> matStr;
        if (chdir(prefix.c_str()) != 0) throw mcg::Error(MCG_EINVAL, "Invalid directory");
        double T = 300.;
        argss >> T;
        std::unique_ptr<Material> mat;
        if (matStr == "grey") {
            mat.reset(new Material("grey_disp.txt", "grey_relax2.txt", T));
        } else if (matStr == "silicon") {
            mat.reset(new Material("Si_disp.txt", "Si_relax2.txt", T));
        } else if (matStr == "custom") {
            std::string disp, relax;
            argss >> disp >> relax;
            mat.reset(new Material(disp, relax, T));
        } else {
            throw mcg::Error(MCG_EINVAL, "Invalid material");
        }
        if (argss.fail()) throw mcg::Error(MCG_EINVAL, "Invalid material arguments");
        std::cout << *mat << "\n\n";

        argss >> domStr;
        std::unique_ptr<Domain> dom;
        if (domStr == "bulk") {
            double d0;
            long div0;
            argss >> d0 >> div0;
            dom.reset(new BulkDomain(Vector3d(d0, d0, d0), Vector3l(div0, 0, 0), 1e6 * d0));
        } else if (domStr == "film") {
            double d0, d1;
            long div1;
            argss >> d0 >> d1 >> div1;
            dom.reset(new FilmDomain(Vector3d(d0, d1, d0), Vector3l(0, div1, 0), 1e6 * d0));
        } else if (domStr == "slab" || domStr == "bar") {
            double L, W, th, tc;
            long div;
            argss >> L >> W >> div >> th >> tc;
            if (domStr == "slab")
                dom.reset(new SlabDomain(Vector3d(L, W, W), Vector3l(div, 0, 0), th, tc));
            else
                dom.reset(new BarDomain(Vector3d(L, W, W), Vector3l(div, 0, 0), th, tc));
        } else if (domStr == "jct") {
            // main.cpp:339-349: dim = (d0, d0, d0, d3), no divisions, dT = 2e6 d0
            double d0, d3;
            argss >> d0 >> d3;
            dom.reset(new JctDomain({d0, d0, d0, d3}, {0, 0, 0, 0}, 2e6 * d0));
        } else if (domStr == "tee") {
            // main.cpp:350-363: dim = (d0 x4, d4), div = (n x4, 0), dT = 3e6 d0
            double d0, d4;
            long n;
            argss >> d0 >> d4 >> n;
            dom.reset(new TeeDomain({d0, d0, d0, d0, d4}, {n, n, n, n, 0}, 3e6 * d0));
        } else if (domStr == "tube") {
            // main.cpp:364-375: dim = (d0, d1, d1, d3), div = (0, n1, n1, n3), dT = 1e6 d0
            double d0, d1, d3;
            long n1, n3;
            argss >> d0 >> d1 >> d3 >> n1 >> n3;
            dom.reset(new TubeDomain({d0, d1, d1, d3}, {0, n1, n1, n3}, 1e6 * d0));
        } else if (domStr == "octet") {
            // main.cpp:376-386: dim0..dim3, div0..div3, dT
            std::array<double, 4> dim{};
            std::array<long, 4> div{};
            double dT = 0.;
            for (double& v : dim) argss >> v;
            for (long& v : div) argss >> v;
            argss >> dT;
            if (!argss.fail())
                for (double v : dim)
                    if (!(v > 0.)) throw mcg::Error(MCG_EINVAL, "Dimensions must be positive");
            if (!argss.fail()) dom.reset(new OctetDomain(dim, div, dT));
        } else if (domStr == "porous") {
            double side, thick, radius;
            long div;
            argss >> side >> thick >> radius >> div;
            dom.reset(new PorousFilmDomain<Diff>(Vector3d(side, side, thick), radius, Vector3l(div, div, 0), 1.0));
        } else {
            throw mcg::Error(MCG_EINVAL, "Invalid domain");
        }
        if (argss.fail()) throw mcg::Error(MCG_EINVAL, "Invalid domain arguments");
        std::cout << *dom << "\n\n";
        if (std::getenv("MCG_DUMP_SCENE")) {
            dump_scene(dom->spec());
            return 0;
        }

        argss >> probStr;
        long nemit = 0, size = 0, maxscat = 0, maxloop = 0, nsim = 0;
        if (probStr == "check") {
            // main.cpp:407-414 and checkDomain (main.cpp:104-143): from every
            // check point, a two-scatter trajectory along each column of rot
            // and its opposite; then all of them, NaN-separated
            Matrix3d rot;
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 3; ++c) argss >> rot(r, c);
            if (argss.fail()) throw mcg::Error(MCG_EINVAL, "Invalid problem arguments");
            const Matrix3Xd pts = dom->checkpoints();
            const char* se = std::getenv("MCG_SEED");
            const uint64_t seed = se ? std::strtoull(se, nullptr, 10) : uint64_t(std::random_device()());
            Rng gen(seed);
            std::vector<ArrayXXd> trajs;
            long total = 0, n = 0;
            for (long q = 0; q < pts.cols(); ++q) {
                const Vector3d pos = pts.col(int(q));
                for (int i = 0; i < 6; ++i) {
                    Vector3d dir = rot.col(i % 3);
                    if (i >= 3) dir = dir * -1.0;
                    TrajProblem tp(mat.get(), dom.get(), pos, dir, 2, 2);
                    std::cout << tp << "\n\n";
                    std::cout << "Trajectory " << n++ << "\n  seeds: " << seed << "\n";
                    trajs.push_back(tp.solve(gen, nullptr));
                    const std::vector<double>& h = tp.hops();
                    const size_t nh = h.size() / MCG_TRACE_WIDTH;
                    if (nh > 0 && h[(nh - 1) * MCG_TRACE_WIDTH + 13] == double(MCG_ESCAPED)) std::cout << "Escaped\n";
                    printSolution(trajs.back());
                    total += trajs.back().cols() + 1;
                }
            }
            ArrayXXd all(3, std::max<long>(0, total - 1));
            for (long i = 0; i < all.size(); ++i) all.data()[i] = std::numeric_limits<double>::quiet_NaN();
            long j = 0;
            for (const ArrayXXd& t : trajs) {
                for (long c = 0; c < t.cols(); ++c)
                    for (int r = 0; r < 3; ++r) all(r, j + c) = t(r, c);
                j += t.cols() + 1;
            }
            std::cout << "Combined Trajectory\n";
            printSolution(all);
            return 0;
        }
        if (probStr == "traj") {
            // main.cpp:416-426: pos, dir, maxscat, maxloop; solveTraj (main.cpp:86-102)
            Vector3d pos, dir;
            argss >> pos(0) >> pos(1) >> pos(2) >> dir(0) >> dir(1) >> dir(2) >> maxscat >> maxloop;
            if (argss.fail()) throw mcg::Error(MCG_EINVAL, "Invalid problem arguments");
            TrajProblem tp(mat.get(), dom.get(), pos, dir, maxscat, maxloop);
            const char* pr = std::getenv("MCG_PRECISION");
            tp.precision(pr && std::string(pr) == "fp32" ? MCG_FP32_FLIGHT : MCG_FP64);
            std::cout << tp << "\n\n";
            const char* se = std::getenv("MCG_SEED");
            const uint64_t seed = se ? std::strtoull(se, nullptr, 10) : uint64_t(std::random_device()());
            Rng gen(seed);
            std::cout << "Trajectory 0\n  seeds: " << seed << "\n";
            ArrayXXd traj = tp.solve(gen, nullptr);
            const std::vector<double>& h = tp.hops();
            const size_t nh = h.size() / MCG_TRACE_WIDTH;
            if (nh > 0 && h[(nh - 1) * MCG_TRACE_WIDTH + 13] == double(MCG_ESCAPED)) std::cout << "Escaped\n";
            printSolution(traj);
            std::cout << "  key: " << tp.lastKey() << "\n";
            return 0;
        }
        std::unique_ptr<FieldProblem> prob;
        if (probStr == "temp" || probStr == "flux" || probStr == "multi") {
            argss >> nemit >> maxscat >> maxloop >> nsim;
            if (probStr == "temp") prob.reset(new TempProblem(mat.get(), dom.get(), nemit, maxscat, maxloop));
            if (probStr == "flux") prob.reset(new FluxProblem(mat.get(), dom.get(), nemit, maxscat, maxloop));
            if (probStr == "multi") prob.reset(new MultiProblem(mat.get(), dom.get(), nemit, maxscat, maxloop));
        } else if (probStr == "cumtemp" || probStr == "cumflux") {
            argss >> nemit >> size >> maxscat >> maxloop >> nsim;
            if (probStr == "cumtemp")
                prob.reset(new CumTempProblem(mat.get(), dom.get(), nemit, size, maxscat, maxloop));
            else
                prob.reset(new CumFluxProblem(mat.get(), dom.get(), nemit, size, maxscat, maxloop));
        } else {
            throw mcg::Error(MCG_EINVAL, "Invalid problem");
        }
        if (argss.fail()) throw mcg::Error(MCG_EINVAL, "Invalid problem arguments");
        const char* pr = std::getenv("MCG_PRECISION");
        prob->precision(pr && std::string(pr) == "fp64" ? MCG_FP64 : MCG_FP32_FLIGHT);
        const mcg_problem_info& info = prob->plan();
        std::cout << "FieldProblem " << prob.get() << "\n  nemit:   " << info.nemit
                  << "\n  maxscat: " << maxscat << "\n  maxloop: " << info.maxloop
                  << "\n  power:   " << info.power << "\n\n";

        const char* se = std::getenv("MCG_SEED");
        uint64_t seed = se ? std::strtoull(se, nullptr, 10) : uint64_t(std::random_device()());
        Rng gen(seed);
        auto t0 = std::chrono::steady_clock::now();
        std::cout << "  seeds: " << seed << "\n";
        if (nsim <= 1) {
            solveField(prob.get(), gen, &std::cout);
            const mcg_stats& st = prob->lastStats();
            std::cout << "  steps: " << st.steps << "  esc: " << st.escapes << "  kernel_ms: " << st.kernel_ms
                      << "\n";
        } else {
            calcStats(prob.get(), nsim, gen, &std::cout);
        }
        double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        std::cout << "Total time: " << secs << " s\n\n";
    } catch (const mcg::Error& e) {
        std::cerr << "error: " << e.what() << "\n";
        return 1;
    }
    return 0;
}
