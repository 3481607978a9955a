// mcg_scene.h -- scene layouts shared by the host builder and the HIP kernels.
//
// The host derives everything the reference's constructors compute (plane
// normals and offsets, rotation and reflection matrices, periodic maps, cell
// inverse matrices and tolerances, tally strides, emitter registry, sampling
// tables) once, in FP64, and uploads it in two precisions.  Kernels read these
// flat records; nothing on the device re-derives geometry.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/mcg.h"

namespace mcg {

// Word slots of one Philox event (event 0 = emission, event i + 1 = the
// scatter after loop iteration i).  A double uniform in slot w takes its high
// word from w and its low word from w + 4, the same word of the next block;
// the FP32 flight path uses the high word alone, so a scatter costs it one
// Philox block (block 1 only for a partial-specularity test).  Free-path
// redraws use adjacent pairs (W_REDRAW + 2k, + 2k + 1).
enum : int {
    W_DIR1 = 0,
    W_DIR2 = 1,
    W_PATH = 2,
    W_MODE = 3,
    W_SPEC = 7,
    W_POS1 = 8,
    W_POS2 = 9,
    W_POS3 = 10,
    W_PIECE = 11,
    W_REDRAW = 16
};

enum ShapeKind : int32_t { SHAPE_PAR = 0, SHAPE_TRI = 1, SHAPE_POLY = 2 };

// Triangles of a polygon face's fan, or pieces of a prism / pyramid's volume
constexpr int kMaxPieces = MCG_MAX_SIDES - 2;

template <typename F>
struct DevPlane {
    F n[3];
    F c;              // signed distance = n . p + c
    F rot[9];         // column-major rotation taking z onto n
    F refl[9];        // column-major I - 2 n n^T
    F prot[9];        // periodic map rotation (column-major)
    F pt[3];          // periodic map translation
    int32_t kind;     // mcg_bdry_kind
    int32_t sdom;     // owning subdomain
    int32_t target;   // Peri: subdomain of the paired face; Inter with one partner: that partner
    int32_t ibeg;     // Inter: first candidate in the pair list
    int32_t icnt;     // Inter: number of candidates
    uint32_t spec_thr;
};

template <typename F>
struct DevSdom {
    F o[3];
    F inv[9];         // column-major inverse of the edge matrix
    F divf[3];        // tally divisions as reals (coord = div * inv (p - o))
    F eps;            // kill tolerance (100 eps |M|)
    int32_t max[3];   // largest cell index per axis
    int32_t accum;    // -2 none, -1 single cell, 0..2 1-D axis, 3/4 2-D/3-D
    int32_t pbeg, pend;
    int32_t col_off;
    int32_t stride[3];
    int32_t pore;     // index of the cylindrical pore or -1
    int32_t box;      // 1: six axis-aligned faces (the FP32 flight's box advect)
    int32_t gax[2];   // 2-D grid (accum 3): its two divided axes
    F bmin[3], bmax[3];        // box: the faces' positions on each axis
    F kmin[3], kmax[3];        // box: positions beyond which a bundle is lost (eps outside)
    int32_t fmin[3], fmax[3];  // box: plane offsets (from pbeg) of the low / high face
};

template <typename F>
struct DevEmitter {
    F o[3];
    F si[3], sj[3];   // wall shape vectors
    F mat[9];         // volume: edge matrix (prism / pyramid: column 2 only)
    F rot[9];         // wall: rotMatrix(normal); volume: rotMatrix(gradT)
    F grad[3];
    int32_t volume;
    int32_t shape;    // wall: ShapeKind; volume: mcg_cell_kind
    int32_t sdom;
    int32_t sign;     // wall: T >= 0
    // SHAPE_POLY walls and prism / pyramid volumes: piece k is drawn when
    // u < pcum[k] for the first such k (pcum is cumulative, normalised, and
    // exactly 1 from the last piece of nonzero weight on); a polygon's piece
    // is the triangle (pa, pb), a volume's the triangular prism or
    // tetrahedron with edges (pa, pb, mat column 2)
    int32_t npiece;
    F pa[kMaxPieces][3], pb[kMaxPieces][3];
    F pcum[kMaxPieces];
};

template <typename F>
struct DevPore {
    F cx, cy, r, r2;
    F eps;
    int32_t kind;
    uint32_t spec_thr;
};

// Everything the host derives for one (material, domain) pair, FP64.
struct HostScene {
    // material
    int64_t nw = 0, np = 0;
    double temp = 0;
    double energy_sum = 0, flux_sum = 0, scat_sum = 0, cond = 0;
    std::vector<double> tau, vel, energy, flux, scat;   // nw x np
    std::vector<int64_t> mode_cell;
    std::vector<double> mode_vel, mode_mfp;
    std::vector<uint64_t> flux_alias, scat_alias;       // thr | alias << 32
    // geometry
    std::vector<DevPlane<double>> planes;
    std::vector<DevSdom<double>> sdoms;
    std::vector<DevEmitter<double>> emitters;
    std::vector<DevPore<double>> pores;
    std::vector<double> emit_weight;
    std::vector<int32_t> inter_list;      // candidate subdomains of Inter faces
    std::vector<double> cell_vol;         // per tally column
    int64_t cols = 0;
    bool has_1d = false;     // some subdomain tallies along one axis (range adds)
    bool has_grid = false;   // some subdomain tallies on a 2-D / 3-D grid
    bool grid_2d = false;    // no grid is 3-D (with has_grid: every grid is 2-D, accum 3)
    int64_t grid_cells = 0;  // cells of the largest 2-D / 3-D grid
};

struct HostProblem {
    int tally = MCG_MULTI;
    int precision = MCG_FP64;
    int64_t rows = 0;
    int64_t nemit = 0, maxscat = 0, maxloop = 0, size = 0, step = 0;
    double power = 0, weight_sum = 0;
    std::vector<int64_t> emit_count, emit_cdf;
    uint64_t seed = 0;
    int64_t first = 0, count = 0;
};

// host builders (mcg_host.cpp); each returns an mcg_status and sets the
// thread's last error message on failure
int build_material(const mcg_material_spec* spec, HostScene& s);
int build_domain(const mcg_domain_spec* spec, HostScene& s);
// trace_only: a trajectory from a given start (TrajProblem with pos_ set,
// problem.cpp:226-299) draws nothing from the emitters, so the domain needs
// no source; nemit, counts and power are then left empty
int build_problem(const mcg_problem_spec* spec, const HostScene& s, HostProblem& p, bool trace_only = false);
int build_alias(const double* pdf, int64_t n, std::vector<uint64_t>& table);
// out (rows x cols, column-major) = postProc of the raw sums, times power_
// over the cell volumes (problem.cpp:439-444): direct is rows x cols and diff
// rows x (cols + 1) (range adds, or NULL), both row-major
void finalise_field(const HostScene& s, const HostProblem& p, const double* direct,
                    const double* diff, double* out);
// the same from the folded raw sums already in sol (column-major), in place
void normalise_field(const HostScene& s, const HostProblem& p, double* sol);

int set_error(int code, const std::string& msg);

}  // namespace mcg
