// mcg_kernels.hip -- the trajectory kernel and the scene / solve entry points.
//
// One launch runs a whole field solve (FieldProblem::solve, problem.cpp:370-445)
// for a bundle range.  The kernel is persistent and history-based: every lane
// owns one live bundle in registers and, as soon as that bundle terminates,
// takes the next bundle id from its wavefront's pool (refilled 256 ids at a
// time with one global atomic), so lanes stay busy until the bundle set is
// exhausted.  Each bundle's random numbers depend only on (seed, bundle id,
// event index), which makes results independent of scheduling.
//
// Per workgroup, the scatter sampling table and the per-mode velocity and mean
// free path live in LDS, as do the subdomain planes and, when small enough,
// the tally: FP64 LDS atomics (ds_add_f64), with the interior of 1-D flights
// added as a range through a difference array (two atomics instead of one per
// crossed cell), flushed once per workgroup with global FP64 atomics.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <type_traits>
#include <vector>

#include "mcg_device.h"
#include "mcg_scene.h"

// Device memory of mcg_solve_static, the solve that allocates nothing: every
// per-solve buffer of that entry point lives here.  Each GPU the library runs
// on holds its own copy until the process ends.  In host code this name is a
// stand-in, not a device address: the host side takes each device's address
// with hipGetSymbolAddress once (static_scratch below) and uses only that.
static __device__ __align__(256) unsigned char scratch[MCG_STATIC_SCRATCH_BYTES];

namespace mcg {

namespace {

constexpr int kWave = 64;
constexpr int kChunk = 256;            // bundle ids claimed per wavefront grab
constexpr int kMaxRows = 4;            // nonzero amount entries per step
constexpr int kMaxFaces = 6;           // faces an FP32 step keeps in registers (the parallelepiped's)
constexpr size_t kLdsTallyCap = 48 * 1024;
// Emission inside a pored box: position redraw k uses Philox blocks
// kPosRedrawBlock + 2k (two coordinates) and + 2k + 1 (the third); a pore
// covers at most pi/4 of the cross-section, so the cap is never reached in
// practice (probability below 1e-400).
constexpr uint32_t kPosRedrawBlock = 64;
constexpr uint32_t kMaxPosRedraw = 4096;

// Scenes with 2-D / 3-D tallies defer the tally walks (kernels with
// kDefer): a wavefront's lanes cross very different numbers of
// cells per step (14 on average, 56 for the longest of 64 lanes on the
// 64 x 64 porous-film map), so walking them to completion inside the step
// leaves most lanes idle.  Deferred, a lane walks in the wavefront's walk
// phase, which runs while at least kWalkQuorum lanes still walk; lanes whose
// walks are done take their next step meanwhile.  Grids of a few cells walk
// at once (kDeferMinCells), and so does the FP64 batch kernel, which would
// spill the walk state: the same additions, without the walk state's
// registers held across steps.
// Lanes whose bundle ended claim and emit new bundles together: emission is
// a long stretch of code (three Philox blocks, the emitter search, the flux
// alias draw, direction and free path), and run whenever any one lane of a
// wavefront needs it (a bar bundle ends after 27 steps, so almost every
// iteration) it would execute with two or three lanes active.  A wavefront
// emits once at least kEmitQuorum lanes wait, or when no lane has anything
// else to do.
#ifndef MCG_EMIT_QUORUM
#define MCG_EMIT_QUORUM 8
#endif
constexpr int kEmitQuorum = MCG_EMIT_QUORUM;
// FP64 single solve: its emission is longer (four Philox blocks and the
// portable transcendentals), so its lanes wait for more peers
#ifndef MCG_EMIT_QUORUM_FP64
#define MCG_EMIT_QUORUM_FP64 16
#endif
constexpr int kEmitQuorumFp64 = MCG_EMIT_QUORUM_FP64;
// FP32 single solve: empty next-bundle slots a wavefront collects before it
// emits into them
#ifndef MCG_SLOT_QUORUM
#define MCG_SLOT_QUORUM 16
#endif
constexpr int kSlotQuorum = MCG_SLOT_QUORUM;
// FP32 single solve: the end phase starts this many bundles per lane of the
// grid before the last id
#ifndef MCG_TAIL_LANES
#define MCG_TAIL_LANES 2
#endif

#ifndef MCG_WALK_QUORUM
#define MCG_WALK_QUORUM 20
#endif
constexpr int kWalkQuorum = MCG_WALK_QUORUM;
// crossings a walking lane advances between checks of the quorum
#ifndef MCG_WALK_UNROLL
#define MCG_WALK_UNROLL 4
#endif
constexpr int kWalkUnroll = MCG_WALK_UNROLL;

enum Counter : int { C_NEXT = 0, C_STEPS, C_ESC, C_ABS, C_TRUNC, C_BUNDLES, C_COUNT };

template <typename F>
struct KParams {
    const DevPlane<F>* planes;
    const DevSdom<F>* sdoms;
    const DevEmitter<F>* emitters;
    const DevPore<F>* pores;
    const int64_t* emit_cdf;
    const int32_t* inter_list;
    const F* mode_vel;
    const F* mode_mfp;
    const uint64_t* scat_alias;
    const uint64_t* flux_alias;
    int32_t nplanes, nsdom, nemitter, npore, nmodes;
    int32_t tally, rows, cols;
    int64_t step_cum, maxscat;
    uint32_t maxloop;
    uint32_t k0, k1;
    int64_t first, count;
    double* direct;   // rows x cols, row-major
    double* diff;     // rows x (cols + 1), row-major
    unsigned long long* counters;
    unsigned long long* next;   // bundle claim counter (offset into [first, first + count))
    uint32_t* c_nscat;
    uint32_t* c_nbdry;
    uint32_t* c_nstep;
    uint32_t* c_fate;
    int32_t lds_tally;
    int32_t lds_diff;      // the LDS tally holds the range-add difference array (1-D tallies)
    int32_t chunk;         // bundle ids a wavefront claims at a time
    int64_t tail;          // FP32 single solve: ids from the end where claims shrink (0: never)
    int64_t trace_id;      // bundle whose steps are recorded, -1 for none
    double* trace;         // MCG_TRACE_WIDTH doubles per step
    int64_t trace_cap;     // records available
    // a tally in global memory, privatised: workgroup w adds into copy
    // w % ncopies (copy 0 = direct / diff, copy c > 0 = priv + (c - 1) *
    // copy_stride: direct, then diff); ncopies <= 1: one copy
    double* priv;
    int64_t copy_stride;
    int32_t ncopies;
};

// Tally sink: LDS when staged, global otherwise (both FP64 atomic adds).
struct Sink {
    double* direct;
    double* diff;
    int32_t cols;
    __device__ __forceinline__ void add(int row, int col, double v) const {
        unsafeAtomicAdd(direct + row * cols + col, v);
    }
    // add at a flat index row * cols + col
    __device__ __forceinline__ void add_at(int idx, double v) const { unsafeAtomicAdd(direct + idx, v); }
    __device__ __forceinline__ void add_range(int row, int lo, int hi, double v) const {
        // +v on columns lo..hi through the difference array
        unsafeAtomicAdd(diff + row * (cols + 1) + lo, v);
        unsafeAtomicAdd(diff + row * (cols + 1) + hi + 1, -v);
    }
};

template <typename F>
struct Lane {
    dev::V3<F> pos, dir;
    F scat_next;
    uint32_t mode;
    int32_t sdom;
    int32_t sign;
    uint32_t nscat, nbdry, nstep;
    int64_t id;
};

// Whether p is outside the pore's cylinder (in the material), with the
// subdomain's tolerance.
template <typename F>
__device__ __forceinline__ bool outside_pore(const DevPore<F>& Q, const dev::V3<F>& p) {
    const F dx = p.x - Q.cx, dy = p.y - Q.cy;
    const F rr = Q.r - Q.eps;
    return !(dx * dx + dy * dy < rr * rr);
}

template <typename F>
__device__ __forceinline__ bool inside(const DevPlane<F>* planes, const DevSdom<F>& S,
                                       const DevPore<F>* pores, const dev::V3<F>& p) {
    for (int b = S.pbeg; b < S.pend; ++b) {
        const DevPlane<F>& P = planes[b];
        if (dev::dot3(P.n, p) + P.c < -S.eps) return false;
    }
    return S.pore < 0 || outside_pore(pores[S.pore], p);
}

// Block j of an event when it holds only the low words of uniforms whose
// high words are in block j - 1: FP64 evaluates it, FP32 (which reads high
// words alone) does not.
template <typename F>
__device__ __forceinline__ uint4 low_block(const dev::Event& ev, uint32_t j) {
    if constexpr (sizeof(F) == 8) return ev.block(j);
    else return make_uint4(0u, 0u, 0u, 0u);
}

// The partial-specularity word (word 3 of block 1): FP64 has the block at
// hand, FP32 evaluates it here, on a partial face only.
template <typename F>
__device__ __forceinline__ uint32_t spec_word(const dev::Event& ev, const uint4& b1) {
    if constexpr (sizeof(F) == 8) return b1.w;
    else return ev.block(1).w;
}

// Free path: l = v tau (-ln(1 - u)), redrawn while l < the smallest normal
// (material.cpp:215-224).
template <typename F>
__device__ __forceinline__ F draw_path(const dev::Event& ev, const uint4& h, const uint4& l, F mfp) {
    F d = mfp * dev::Real<F>::neglog1m(dev::Real<F>::uniform(h.z, l.z));
    for (uint32_t k = 0; d < dev::Real<F>::kMin; ++k) {
        const uint4 b = ev.block(4u + k / 2u);
        const uint32_t a = (k & 1u) ? b.z : b.x, c = (k & 1u) ? b.w : b.y;
        d = mfp * dev::Real<F>::neglog1m(dev::Real<F>::uniform(a, c));
    }
    return d;
}

// The piece of a polygon face or N-gon cell a uniform selects: the first k
// with u < pcum[k] (pcum reaches exactly 1 at the last piece drawn).
template <typename F>
__device__ __forceinline__ int pick_piece(const DevEmitter<F>& E, F u) {
    int k = 0;
    while (k + 1 < E.npiece && !(u < E.pcum[k])) ++k;
    return k;
}

// Emitter::emit with the wall / volume draws, then drawScatNext
// (problem.cpp:386-399, boundary.cpp:378-431 and 244-251 for a polygon,
// subdomain.cpp:255-441).
template <typename F>
__device__ void emit(const KParams<F>& K, const uint64_t* flux_alias, const F* mfp_tab,
                     int64_t n, Lane<F>& L) {
    int lo = 0, hi = K.nemitter;   // upper_bound of n in the cumulative counts
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (K.emit_cdf[mid] <= n) lo = mid + 1; else hi = mid;
    }
    const DevEmitter<F>& E = K.emitters[lo];
    const dev::Event ev{K.k0, K.k1, 0u, uint32_t(n), uint32_t(uint64_t(n) >> 32)};
    const uint4 b0 = ev.block(0), b1 = low_block<F>(ev, 1), b2 = ev.block(2), b3 = low_block<F>(ev, 3);
    L.mode = dev::alias_draw(flux_alias, uint32_t(K.nmodes), b0.w);
    dev::V3<F> d;
    if (!E.volume) {
        const F r1 = dev::Real<F>::uniform(b2.x, b3.x), r2 = dev::Real<F>::uniform(b2.y, b3.y);
        F a = r1, b = r2;
        if (E.shape != SHAPE_PAR && !(r1 + r2 < F(1))) {
            a = F(1) - r1;
            b = F(1) - r2;
        }
        F si[3] = {E.si[0], E.si[1], E.si[2]}, sj[3] = {E.sj[0], E.sj[1], E.sj[2]};
        if (E.npiece) {   // a polygon: one triangle of its fan, by area
            const int k = pick_piece(E, dev::Real<F>::uniform(b2.w, b3.w));
            for (int r = 0; r < 3; ++r) {
                si[r] = E.pa[k][r];
                sj[r] = E.pb[k][r];
            }
        }
        L.pos = dev::V3<F>{E.o[0] + (a * si[0] + b * sj[0]), E.o[1] + (a * si[1] + b * sj[1]),
                           E.o[2] + (a * si[2] + b * sj[2])};
        const int pk = K.npore > 0 ? K.sdoms[E.sdom].pore : -1;
        if (pk >= 0) {   // a wall across the pore (a box face): redraw inside its disk
            const DevPore<F>& Q = K.pores[pk];
            for (uint32_t k = 0; k < kMaxPosRedraw; ++k) {
                const F dx = L.pos.x - Q.cx, dy = L.pos.y - Q.cy;
                if (!(dx * dx + dy * dy < Q.r2)) break;
                const uint4 ba = ev.block(kPosRedrawBlock + 2u * k), bb = low_block<F>(ev, kPosRedrawBlock + 2u * k + 1u);
                const F s1 = dev::Real<F>::uniform(ba.x, bb.x), s2 = dev::Real<F>::uniform(ba.y, bb.y);
                L.pos = dev::V3<F>{E.o[0] + (s1 * E.si[0] + s2 * E.sj[0]), E.o[1] + (s1 * E.si[1] + s2 * E.sj[1]),
                                   E.o[2] + (s1 * E.si[2] + s2 * E.sj[2])};
            }
        }
        d = dev::matvec(E.rot, dev::draw_aniso<F>(b0, b1, false));
        L.sign = E.sign ? 1 : -1;
    } else {
        F c0 = dev::Real<F>::uniform(b2.x, b3.x), c1 = dev::Real<F>::uniform(b2.y, b3.y),
          c2 = dev::Real<F>::uniform(b2.z, b3.z);
        if (E.shape != MCG_PARALLELEPIPED && c0 + c1 > F(1)) {
            c0 = F(1) - c0;
            c1 = F(1) - c1;
        }
        if (E.shape == MCG_TETRAHEDRON || E.shape == MCG_PYRAMID) {
            if (c1 + c2 > F(1)) {
                const F t = c2;
                c2 = F(1) - c0 - c1;
                c1 = F(1) - t;
            } else if (c0 + c1 + c2 > F(1)) {
                const F t = c2;
                c2 = c0 + c1 + c2 - F(1);
                c0 = F(1) - c1 - t;
            }
        }
        F M[9];
        for (int r = 0; r < 9; ++r) M[r] = E.mat[r];
        if (E.npiece) {   // a prism / pyramid: one piece, by its volume term
            const int k = pick_piece(E, dev::Real<F>::uniform(b2.w, b3.w));
            for (int r = 0; r < 3; ++r) {
                M[r] = E.pa[k][r];
                M[3 + r] = E.pb[k][r];
            }
        }
        const dev::V3<F> m = dev::matvec(M, dev::V3<F>{c0, c1, c2});
        L.pos = dev::V3<F>{E.o[0] + m.x, E.o[1] + m.y, E.o[2] + m.z};
        const int pk = K.npore > 0 ? K.sdoms[E.sdom].pore : -1;
        if (pk >= 0) {   // uniform in the box minus the pore: redraw inside it
            const DevPore<F>& Q = K.pores[pk];
            for (uint32_t k = 0; k < kMaxPosRedraw; ++k) {
                const F dx = L.pos.x - Q.cx, dy = L.pos.y - Q.cy;
                if (!(dx * dx + dy * dy < Q.r2)) break;
                const uint4 ba = ev.block(kPosRedrawBlock + 2u * k), bb = low_block<F>(ev, kPosRedrawBlock + 2u * k + 1u);
                const dev::V3<F> q = dev::matvec(
                    E.mat, dev::V3<F>{dev::Real<F>::uniform(ba.x, bb.x), dev::Real<F>::uniform(ba.y, bb.y),
                                      dev::Real<F>::uniform(ba.z, bb.z)});
                L.pos = dev::V3<F>{E.o[0] + q.x, E.o[1] + q.y, E.o[2] + q.z};
            }
        }
        d = dev::matvec(E.rot, dev::draw_aniso<F>(b0, b1, true));
        L.sign = (d.x * E.grad[0] + d.y * E.grad[1] + d.z * E.grad[2] < F(0)) ? 1 : -1;
    }
    L.dir = dev::normalized(d);
    L.sdom = E.sdom;
    L.scat_next = draw_path<F>(ev, b0, b1, mfp_tab[L.mode]);
    L.nscat = L.nbdry = L.nstep = 0;
    L.id = n;
}

// A 2-D / 3-D tally walk in progress: each axis's next crossing parameter
// pk (2 once the axis has none left), the plane index of that crossing and
// of the axis's last one (FP32: nxt counts the crossings left and last is
// unused), the column step of each axis, the current column, the parameter
// reached and the step's amounts (exact in F: a sign times an F value).  q is
// the segment's coordinate change (FP64, see crossing_param) or the
// parameter step |1/dc| between an axis's crossings (FP32).
template <typename F>
struct Walk {
    F bc[3], q[3], pk[3];
    int32_t nxt[3], last[3], dcol[3];
    int32_t col, base, nv;   // base: the first row's flat offset, r0 * cols
    F prev;
    F amt[kMaxRows];
};

// Parameter along a segment of the crossing of grid plane n on one axis:
// (n - b) / d.  FP64 divides, as the checker does (deterministic mode is
// compared bit for bit); FP32 flight multiplies by the segment's reciprocal,
// a rounding difference well inside its statistical comparison.
template <typename F>
__device__ __forceinline__ F crossing_param(int n, F b, F q) {
    if constexpr (sizeof(F) == 8) return (F(n) - b) / q;
    else return (F(n) - b) * q;
}

// The crossings of the ordered map of field.cpp:156-218, one call per map
// entry: crossings of several axes at exactly the same parameter are one
// entry.  Adds the segment up to the next crossing to the current cell and
// moves on; returns true after the last segment (up to parameter 1).
template <typename F>
__device__ __forceinline__ bool walk_next(Walk<F>& W, const Sink& sink) {
    const F pmin = fmin(fmin(W.pk[0], W.pk[1]), W.pk[2]);
    const bool done = pmin == F(2);
    const F upto = done ? F(1) : pmin;
    if (!done || W.prev != F(1)) {
        const int idx = W.base + W.col;
        if constexpr (sizeof(F) == 8) {
            const double seg = double(upto - W.prev);
#pragma unroll
            for (int j = 0; j < kMaxRows; ++j)
                if (j < W.nv) sink.add_at(idx + j * sink.cols, double(W.amt[j]) * seg);
        } else {
            // FP32 flight: the share in FP32, accumulated in FP64
            const F seg = upto - W.prev;
#pragma unroll
            for (int j = 0; j < kMaxRows; ++j)
                if (j < W.nv) sink.add_at(idx + j * sink.cols, double(W.amt[j] * seg));
        }
    }
    if (done) return true;
    W.prev = pmin;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        if (W.pk[k] != pmin) continue;
        W.col += W.dcol[k];
        if constexpr (sizeof(F) == 4) {
            W.pk[k] = W.nxt[k] == 0 ? F(2) : W.pk[k] + W.q[k];
            W.nxt[k] -= W.nxt[k] > 0 ? 1 : 0;
        } else if (W.nxt[k] == W.last[k]) {
            W.pk[k] = F(2);
        } else {
            W.nxt[k] += W.dcol[k] > 0 ? 1 : -1;
            W.pk[k] = crossing_param<F>(W.nxt[k], W.bc[k], W.q[k]);
        }
    }
    return false;
}

// The deferred 2-D / 3-D walk (NA = 2 grid axes when every grid of the scene
// is 2-D, else 3): the crossings taken one per call, the axis whose next
// crossing comes first advancing (an axis with none left holds parameter 2).
// An axis's crossings after its first are reached by increments of |1/dc|.
// Crossings of two axes at the same parameter are taken one after the other,
// the first with a zero-length share, which adds nothing: the cells and
// shares are those of the merged map entries of field.cpp:156-218 up to the
// rounding of the increments (FP64: well inside the 1e-5 tally comparison).
// The amounts are FP64 (exact in FP32 flight: a sign times an FP32 value),
// so a share costs one FP64 multiply; the row stride and row count ride in
// the walk's registers, so the loop reads no kernel parameter.
template <typename F, int NA>
struct WalkN {
    F t[NA];                 // next crossing parameter per axis, 2 once none is left
    F dt[NA];                // parameter step between the axis's crossings
    int32_t n[NA];           // crossings left per axis
    int32_t dcol[NA];        // column step of a crossing
    int32_t idx;             // flat index of the current cell's first row
    int32_t rstride;         // row stride (the tally's columns)
    int32_t nv;              // amounts (rows) of the step
    F prev;                  // parameter reached
    double amt[kMaxRows];
};

// One map entry of the deferred walk; true after the last segment.
template <typename F, int NA>
__device__ __forceinline__ bool walk_next(WalkN<F, NA>& W, const Sink& sink) {
    // the parameters are ordinary numbers: a compare and select, without the
    // NaN canonicalisation fmin costs
    F tm = W.t[0];
#pragma unroll
    for (int q = 1; q < NA; ++q) tm = W.t[q] < tm ? W.t[q] : tm;
    const bool done = tm == F(2);
    if (!done || W.prev != F(1)) {
        const double seg = double((done ? F(1) : tm) - W.prev);
        double* p = sink.direct + W.idx;
#pragma unroll
        for (int j = 0; j < kMaxRows; ++j) {
            if (j < W.nv) unsafeAtomicAdd(p, W.amt[j] * seg);
            p += W.rstride;
        }
    }
    if (done) return true;
    W.prev = tm;
    // the first axis at tm advances; selects, not branches, so a crossing
    // costs no exec-mask changes
    bool open = true;   // no axis advanced yet
#pragma unroll
    for (int q = 0; q < NA; ++q) {
        const bool take = open && W.t[q] == tm;
        open = open && !take;
        const int32_t n = W.n[q] - 1;
        W.idx += take ? W.dcol[q] : 0;
        W.n[q] = take ? n : W.n[q];
        W.t[q] = take ? (n > 0 ? W.t[q] + W.dt[q] : F(2)) : W.t[q];
    }
    return false;
}

template <typename F, int kDefer>
struct WalkSel {
    using type = WalkN<F, kDefer>;
};
template <typename F>
struct WalkSel<F, 0> {
    using type = Walk<F>;
};
template <typename F, int kDefer>
using WalkOf = typename WalkSel<F, kDefer>::type;

// Field::accumulate (field.cpp:92-220) of the amounts `famt` (nv entries on
// rows r0..) over the segment a -> b of subdomain S.  With kDefer a 2-D /
// 3-D walk is only set up in W (the caller runs walk_next later) and the
// function returns true; otherwise everything is added before it returns.
template <typename F, int kDefer>
__device__ __forceinline__ bool accumulate(const DevSdom<F>& S, const Sink& sink,
                                           const dev::V3<F>& a, const dev::V3<F>& b, int r0,
                                           int nv, const F (&famt)[kMaxRows], WalkOf<F, kDefer>& W) {
    double amt[kMaxRows];
#pragma unroll
    for (int j = 0; j < kMaxRows; ++j) amt[j] = double(famt[j]);
    const int flag = S.accum;
    if (flag < -1) return false;
    if (flag < 0) {
#pragma unroll
        for (int j = 0; j < kMaxRows; ++j)
            if (j < nv) sink.add(r0 + j, S.col_off, amt[j]);
        return false;
    }
    const dev::V3<F> la{a.x - S.o[0], a.y - S.o[1], a.z - S.o[2]};
    const dev::V3<F> lb{b.x - S.o[0], b.y - S.o[1], b.z - S.o[2]};
    if (flag < 3) {
        // one axis: only that coordinate, row `flag` of the inverse edge
        // matrix with the same operations as the full product.  The column
        // offset is read once: the LDS atomics below could alias the staged
        // subdomain as far as the compiler knows, and it would re-read it.
        const int col0 = S.col_off;
        const F bcd_f = S.divf[flag] * dev::matrow(S.inv, flag, la);
        const F ecd_f = S.divf[flag] * dev::matrow(S.inv, flag, lb);
        const int bb = min(max(int(floor(bcd_f)), 0), S.max[flag]);
        const int ee = min(max(int(floor(ecd_f)), 0), S.max[flag]);
        const int lo = bb < ee ? bb + 1 : ee + 1;
        const int hi = bb < ee ? ee - 1 : bb - 1;
        if constexpr (sizeof(F) == 8) {
            // the checker's FP64 arithmetic: a segment inside one cell adds
            // its amounts there, a longer one each amount divided by |dc|;
            // both kinds of lane share the adds to the first cell
            const bool one = bb == ee;
            const double bcd = double(bcd_f), ecd = double(ecd_f);
            const double ad = fabs(double(ecd_f - bcd_f));
            const double fb = bb < ee ? double(1 + bb) - bcd : bcd - double(bb);
            const double fe = bb < ee ? ecd - double(ee) : double(1 + ee) - ecd;
#pragma unroll
            for (int j = 0; j < kMaxRows; ++j) {
                if (j >= nv) continue;
                const double ca = amt[j] / ad;
                sink.add(r0 + j, col0 + bb, one ? amt[j] : ca * fb);
                if (!one) sink.add(r0 + j, col0 + ee, ca * fe);
                if (lo <= hi) sink.add_range(r0 + j, col0 + lo, col0 + hi, ca);
            }
        } else {
            // FP32 flight: the per-step shares in FP32 (the amounts are FP32
            // values), accumulated in FP64; a segment inside one cell takes
            // the same path (its whole amount to that cell), so lanes do not
            // split on it
            const bool one = bb == ee;
            const F rad = __builtin_amdgcn_rcpf(fabsf(ecd_f - bcd_f));
            const F fb = bb < ee ? F(1 + bb) - bcd_f : bcd_f - F(bb);
            const F fe = bb < ee ? ecd_f - F(ee) : F(1 + ee) - ecd_f;
#pragma unroll
            for (int j = 0; j < kMaxRows; ++j) {
                if (j >= nv) continue;
                const F ca = famt[j] * rad;
                sink.add(r0 + j, col0 + bb, double(one ? famt[j] : ca * fb));
                if (!one) sink.add(r0 + j, col0 + ee, double(ca * fe));
                if (lo <= hi) sink.add_range(r0 + j, col0 + lo, col0 + hi, double(ca));
            }
        }
        return false;
    }
    const dev::V3<F> ia = dev::matvec(S.inv, la), ib = dev::matvec(S.inv, lb);
    const F bc[3] = {S.divf[0] * ia.x, S.divf[1] * ia.y, S.divf[2] * ia.z};
    const F ec[3] = {S.divf[0] * ib.x, S.divf[1] * ib.y, S.divf[2] * ib.z};
    int bi[3], ei[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const int vb = int(floor(bc[k])), ve = int(floor(ec[k]));
        bi[k] = min(max(vb, 0), S.max[k]);
        ei[k] = min(max(ve, 0), S.max[k]);
    }
    if constexpr (kDefer != 0) {
        // the crossings of each walked axis: first parameter, step, count
        int col = S.col_off;
#pragma unroll
        for (int k = 0; k < 3; ++k) col += bi[k] * S.stride[k];
#pragma unroll
        for (int q = 0; q < kDefer; ++q) {
            const int k = kDefer == 3 ? q : S.gax[q];
            const F b0 = k == 0 ? bc[0] : (k == 1 ? bc[1] : bc[2]);
            const F e0 = k == 0 ? ec[0] : (k == 1 ? ec[1] : ec[2]);
            const int bk = k == 0 ? bi[0] : (k == 1 ? bi[1] : bi[2]);
            const int ek = k == 0 ? ei[0] : (k == 1 ? ei[1] : ei[2]);
            const int sk = k == 0 ? S.stride[0] : (k == 1 ? S.stride[1] : S.stride[2]);
            const F dc = e0 - b0;
            const bool up = bk < ek;
            if constexpr (sizeof(F) == 4) {
                const bool none = fabsf(dc) < FLT_MIN || bk == ek;
                const F r = __builtin_amdgcn_rcpf(dc);
                W.t[q] = none ? F(2) : (F(up ? bk + 1 : bk) - b0) * r;
                W.dt[q] = fabsf(r);
                W.n[q] = none ? 0 : (up ? ek - bk : bk - ek);
            } else {
                // FP64: the first crossing as the checker computes it, later
                // ones by increments of |1/dc| (the shares move by rounding
                // only; a tie of two axes becomes two entries, the first of
                // zero length)
                const bool none = fabs(dc) < DBL_MIN || bk == ek;
                const F dcs = none ? F(1) : dc;
                W.t[q] = none ? F(2) : (F(up ? bk + 1 : bk) - b0) / dcs;
                W.dt[q] = F(1) / fabs(dcs);
                W.n[q] = none ? 0 : (up ? ek - bk : bk - ek);
            }
            W.dcol[q] = up ? sk : -sk;
        }
        W.idx = r0 * sink.cols + col;
        W.rstride = sink.cols;
        W.nv = nv;
        W.prev = F(0);
#pragma unroll
        for (int j = 0; j < kMaxRows; ++j) W.amt[j] = double(famt[j]);
        return true;
    } else {
    // 2-D / 3-D: the crossings in parameter order (walk_next)
    W.col = S.col_off;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const F dc = ec[k] - bc[k];
        W.bc[k] = bc[k];
        W.col += bi[k] * S.stride[k];
        W.nxt[k] = W.last[k] = W.dcol[k] = 0;
        W.pk[k] = F(2);
        W.q[k] = F(0);
        if (fabs(double(dc)) < DBL_MIN || bi[k] == ei[k]) continue;
        if (bi[k] < ei[k]) {
            W.nxt[k] = bi[k] + 1;
            W.last[k] = ei[k];
            W.dcol[k] = S.stride[k];
        } else {
            W.nxt[k] = bi[k];
            W.last[k] = ei[k] + 1;
            W.dcol[k] = -S.stride[k];
        }
        if constexpr (sizeof(F) == 8) W.q[k] = dc;
        else W.q[k] = __builtin_amdgcn_rcpf(dc);   // FP32 flight: the hardware reciprocal
        W.pk[k] = crossing_param<F>(W.nxt[k], bc[k], W.q[k]);
        if constexpr (sizeof(F) == 4) {
            // FP32 flight steps the parameter by |1/dc| per crossing and
            // counts the crossings left in nxt
            W.nxt[k] = W.last[k] > W.nxt[k] ? W.last[k] - W.nxt[k] : W.nxt[k] - W.last[k];
            W.q[k] = fabsf(W.q[k]);
        }
    }
    W.prev = F(0);
    W.base = r0 * sink.cols;
    W.nv = nv;
#pragma unroll
    for (int j = 0; j < kMaxRows; ++j) W.amt[j] = famt[j];
    while (!walk_next<F>(W, sink)) {
    }
    return false;
    }
}

// Scatter off a cylindrical pore wall.  The wall normal at the hit point
// points away from the axis (into the material); the reflection and the
// Lambertian rotation are built from it as for a planar face.
template <typename F>
__device__ __forceinline__ void pore_scatter(const DevPore<F>& Q, const dev::Event& ev, const uint4& b0,
                                             const uint4& b1, Lane<F>& L) {
    // FP64 divides by the radius as the checker does; FP32 flight
    // normalises the offset from the axis directly (the same direction)
    dev::V3<F> n;
    if constexpr (sizeof(F) == 8)
        n = dev::normalized(dev::V3<F>{(L.pos.x - Q.cx) / Q.r, (L.pos.y - Q.cy) / Q.r, F(0)});
    else
        n = dev::normalized(dev::V3<F>{L.pos.x - Q.cx, L.pos.y - Q.cy, F(0)});
    int kind = Q.kind;
    if (kind == MCG_PARTIAL) kind = spec_word<F>(ev, b1) < Q.spec_thr ? MCG_SPEC : MCG_DIFF;
    if (kind == MCG_SPEC) {
        const F nv[3] = {n.x, n.y, n.z};
        F r[9];
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int q = 0; q < 3; ++q) r[3 * c + q] = (q == c ? F(1) : F(0)) - F(2) * (nv[q] * nv[c]);
        L.dir = dev::normalized(dev::matvec(r, L.dir));
        return;
    }
    // rotation taking z onto n (n.z = 0, so never the antiparallel case)
    const F s = dev::Real<F>::sqrt_((F(1) + n.z) * F(2));
    const F invs = F(1) / s;
    const F qx = -n.y * invs, qy = n.x * invs, qz = F(0) * invs, qw = s * F(0.5);
    const F tx = F(2) * qx, ty = F(2) * qy, tz = F(2) * qz;
    const F twx = tx * qw, twy = ty * qw, twz = tz * qw;
    const F txx = tx * qx, txy = ty * qx, txz = tz * qx;
    const F tyy = ty * qy, tyz = tz * qy, tzz = tz * qz;
    const F r[9] = {F(1) - (tyy + tzz), txy + twz, txz - twy,
                    txy - twz, F(1) - (txx + tzz), tyz + twx,
                    txz + twy, tyz - twx, F(1) - (txx + tyy)};
    L.dir = dev::normalized(dev::matvec(r, dev::draw_aniso<F>(b0, b1, false)));
    L.nscat++;
}

// Lanes set in a wavefront mask, as a 32-bit count (scalar instructions;
// a 64-bit count would be compared on the vector unit)
__device__ __forceinline__ int popc(uint64_t m) {
    return __builtin_popcount(uint32_t(m)) + __builtin_popcount(uint32_t(m >> 32));
}

__device__ __forceinline__ unsigned long long wave_sum(unsigned long long v) {
    for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off);
    return v;
}

// A workgroup's view of one solve's scene: LDS copies when staged.
template <typename F>
struct View {
    const DevPlane<F>* planes;
    const DevSdom<F>* sdoms;
    const DevPore<F>* pores;
    const uint64_t* scat;
    const F* vel;
    const F* mfp;
    double* direct;
    double* diff;
};

// Stage one solve's tables, planes, subdomains and pores in LDS and clear
// the LDS tally; the caller synchronises.  Layout: [scat alias | vel | mfp]
// (when kLdsTables), planes, subdomains, pores, tally direct, tally diff
// (only when the scene has 1-D tallies, the only users of range adds).
template <typename F, bool kLdsTables, bool kLdsTally>
__device__ View<F> stage(const KParams<F>& K, unsigned char* lds, int tid, int nthreads) {
    const int ntab = kLdsTables ? K.nmodes : 0;
    View<F> v;
    size_t off = (size_t(ntab) * (8 + 2 * sizeof(F)) + 15) & ~size_t(15);
    DevPlane<F>* s_planes = reinterpret_cast<DevPlane<F>*>(lds + off);
    off = (off + sizeof(DevPlane<F>) * K.nplanes + 15) & ~size_t(15);
    DevSdom<F>* s_sdoms = reinterpret_cast<DevSdom<F>*>(lds + off);
    off = (off + sizeof(DevSdom<F>) * K.nsdom + 15) & ~size_t(15);
    DevPore<F>* s_pores = reinterpret_cast<DevPore<F>*>(lds + off);
    off = (off + sizeof(DevPore<F>) * K.npore + 15) & ~size_t(15);
    double* s_direct = reinterpret_cast<double*>(lds + off);
    double* s_diff = s_direct + size_t(K.rows) * K.cols;
    if (kLdsTables) {
        uint64_t* w_scat = reinterpret_cast<uint64_t*>(lds);
        F* w_vel = reinterpret_cast<F*>(lds + 8 * size_t(ntab));
        F* w_mfp = w_vel + ntab;
        for (int i = tid; i < ntab; i += nthreads) {
            w_scat[i] = K.scat_alias[i];
            w_vel[i] = K.mode_vel[i];
            w_mfp[i] = K.mode_mfp[i];
        }
        v.scat = w_scat;
        v.vel = w_vel;
        v.mfp = w_mfp;
    } else {
        v.scat = K.scat_alias;
        v.vel = K.mode_vel;
        v.mfp = K.mode_mfp;
    }
    const uint32_t* src = reinterpret_cast<const uint32_t*>(K.planes);
    uint32_t* dst = reinterpret_cast<uint32_t*>(s_planes);
    const int np = int(sizeof(DevPlane<F>) / 4) * K.nplanes;
    for (int i = tid; i < np; i += nthreads) dst[i] = src[i];
    src = reinterpret_cast<const uint32_t*>(K.sdoms);
    dst = reinterpret_cast<uint32_t*>(s_sdoms);
    const int ns = int(sizeof(DevSdom<F>) / 4) * K.nsdom;
    for (int i = tid; i < ns; i += nthreads) dst[i] = src[i];
    src = reinterpret_cast<const uint32_t*>(K.pores);
    dst = reinterpret_cast<uint32_t*>(s_pores);
    const int nq = int(sizeof(DevPore<F>) / 4) * K.npore;
    for (int i = tid; i < nq; i += nthreads) dst[i] = src[i];
    if (kLdsTally) {
        const int n = K.rows * K.cols + (K.lds_diff ? K.rows * (K.cols + 1) : 0);
        for (int i = tid; i < n; i += nthreads) s_direct[i] = 0.0;
        // both in LDS, so the compiler addresses them as LDS (ds_add_f64);
        // without 1-D tallies the difference array is neither allocated nor
        // used (range adds come only from 1-D tallies)
        v.direct = s_direct;
        v.diff = s_diff;
    } else {
        // workgroups are dealt to the XCDs in turn, so copy w % ncopies keeps
        // a workgroup's atomics in its own XCD's L2
        const int c = K.ncopies > 1 ? int(blockIdx.x % unsigned(K.ncopies)) : 0;
        v.direct = c == 0 ? K.direct : K.priv + size_t(c - 1) * size_t(K.copy_stride);
        v.diff = c == 0 ? K.diff : v.direct + (K.diff - K.direct);
    }
    v.planes = s_planes;
    v.sdoms = s_sdoms;
    v.pores = s_pores;
    return v;
}

// Add a workgroup's LDS tally into the solve's global tally (after a barrier).
__device__ void flush_tally(const double* s_direct, const double* s_diff, double* g_direct,
                            double* g_diff, int rows, int cols, bool with_diff, int tid, int nthreads) {
    const int nd = rows * cols;
    for (int i = tid; i < nd; i += nthreads) {
        const double v = s_direct[i];
        if (v != 0.0) unsafeAtomicAdd(g_direct + i, v);
    }
    const int nf = with_diff ? rows * (cols + 1) : 0;
    for (int i = tid; i < nf; i += nthreads) {
        const double v = s_diff[i];
        if (v != 0.0) unsafeAtomicAdd(g_diff + i, v);
    }
}

struct LaneStats {
    unsigned long long steps = 0, esc = 0, absorbed = 0, trunc = 0, bundles = 0;
    __device__ void add(uint32_t fate, uint32_t nstep) {
        steps += nstep;
        bundles++;
        esc += fate == MCG_ESCAPED;
        absorbed += fate == MCG_ABSORBED;
        trunc += fate == MCG_MAXSCAT || fate == MCG_MAXLOOP;
    }
    // wave-reduce and add to the solve's counters (all lanes call)
    __device__ void publish(unsigned long long* counters, int lane) {
        const unsigned long long a = wave_sum(steps), b = wave_sum(esc), c = wave_sum(absorbed),
                                 d = wave_sum(trunc), e = wave_sum(bundles);
        if (lane == 0 && e != 0) {
            atomicAdd(counters + C_STEPS, a);
            atomicAdd(counters + C_ESC, b);
            atomicAdd(counters + C_ABS, c);
            atomicAdd(counters + C_TRUNC, d);
            atomicAdd(counters + C_BUNDLES, e);
        }
        steps = esc = absorbed = trunc = bundles = 0;
    }
};

// FP32 flight at an Inter face with several partners, when rounding has
// left the point inside none of them: the partner it is least outside of
// (the deepest face violation, in units of that cell's tolerance), with the
// point moved back onto the violated faces (a shift of the order of the
// rounding), so that the flight continues there instead of the bundle being
// lost.  -1 if no partner is usable.
template <typename F>
__device__ __forceinline__ int inter_nearest(const View<F>& V, const DevPlane<F>& P, const int32_t* list,
                                          dev::V3<F>& pos) {
    int best = -1;
    F best_depth = F(-3.0e38);
    for (int c = 0; c < P.icnt; ++c) {
        const int cand = list[P.ibeg + c];
        const DevSdom<F>& C = V.sdoms[cand];
        F depth = F(3.0e38);
        for (int b = C.pbeg; b < C.pend; ++b) {
            const DevPlane<F>& Q = V.planes[b];
            depth = fminf(depth, (dev::dot3(Q.n, pos) + Q.c) / C.eps);
        }
        if (C.pore < 0 && depth > best_depth) {
            best_depth = depth;
            best = cand;
        }
    }
    if (best >= 0) {
        const DevSdom<F>& C = V.sdoms[best];
        for (int b = C.pbeg; b < C.pend; ++b) {
            const DevPlane<F>& Q = V.planes[b];
            const F d = dev::dot3(Q.n, pos) + Q.c;
            if (d < F(0)) pos = dev::V3<F>{pos.x - d * Q.n[0], pos.y - d * Q.n[1], pos.z - d * Q.n[2]};
        }
    }
    return best;
}

// The intrinsic and planar-face scatters of FP32 flight in one code path:
// the isotropic (material.cpp:226-231) and Lambertian (boundary.cpp:308-312)
// draws share the event's words and differ only in how the polar angle is
// formed, and the new direction of a diffuse, specular or periodic face is
// one product with the face's rotation, reflection or periodic map, so a
// wavefront whose lanes take different events runs the direction code once.
// FP64 keeps the checker's branch-by-branch arithmetic.
template <typename F>
__device__ __forceinline__ void scatter_fp32(const KParams<F>& K, const View<F>& V, int32_t pbeg,
                                             const dev::Event& ev, const uint4& b0, const uint4& b1,
                                             int hit, Lane<F>& L, uint32_t& fate) {
    const bool intr = hit < 0;
    const DevPlane<F>& P = V.planes[intr ? pbeg : hit];
    // the face's kind and its rotation (what a diffuse face, the common
    // case, needs) read together; a specular or periodic face reads its own
    // matrix afterwards
    const int32_t pkind = P.kind;
    F m[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) m[i] = P.rot[i];
    int kind = intr ? -1 : pkind;
    if (kind == MCG_PARTIAL) kind = spec_word<F>(ev, b1) < P.spec_thr ? MCG_SPEC : MCG_DIFF;
    const bool diff = kind == MCG_DIFF;
    // direction: isotropic cos(theta) = r, or Lambertian sin^2(theta) = |r|
    const F r = dev::Real<F>::uniform(b0.x, b1.x) * F(2) + F(-1);
    const F ar = r < F(0) ? -r : r;
    const F sinT = dev::Real<F>::sqrt_(intr ? F(1) - r * r : ar);
    const F cosT = intr ? r : dev::Real<F>::sqrt_(F(1) - ar);
    const F q = dev::Real<F>::uniform(b0.y, b1.y) * F(2) + F(-1);
    F sq, cq;
    dev::Real<F>::sincos_pi(q, sq, cq);
    const dev::V3<F> v{sinT * cq, sinT * sq, cosT};
    const bool turn = diff || kind == MCG_SPEC || kind == MCG_PERI;
    if (kind == MCG_SPEC || kind == MCG_PERI) {
        const F* M = kind == MCG_SPEC ? P.refl : P.prot;
#pragma unroll
        for (int i = 0; i < 9; ++i) m[i] = M[i];
    }
    const dev::V3<F> out = dev::matvec(m, diff ? v : L.dir);
    if (intr || turn) L.dir = dev::normalized(intr ? v : out);
    L.nscat += (intr || diff) ? 1u : 0u;
    L.nbdry += intr ? 0u : 1u;
    if (intr) {
        L.mode = dev::alias_draw(V.scat, uint32_t(K.nmodes), b0.w);
        L.scat_next = draw_path<F>(ev, b0, b1, V.mfp[L.mode]);
    } else if (kind == MCG_PERI) {
        const dev::V3<F> p = dev::matvec(P.prot, L.pos);
        L.pos = dev::V3<F>{p.x + P.pt[0], p.y + P.pt[1], p.z + P.pt[2]};
        L.sdom = P.target;
    } else if (kind == MCG_ISOT) {
        fate = MCG_ABSORBED;
    } else if (kind == MCG_INTER) {
        int target = -1;
        if (P.icnt == 1) {
            target = P.target;
        } else {
            for (int c = 0; c < P.icnt; ++c) {
                const int cand = K.inter_list[P.ibeg + c];
                if (inside(V.planes, V.sdoms[cand], V.pores, L.pos)) {
                    target = cand;
                    break;
                }
            }
            if (target < 0) target = inter_nearest(V, P, K.inter_list, L.pos);
        }
        if (target < 0) fate = MCG_ESCAPED;
        else L.sdom = target;
    }
}

// The flight half of a scatter-step (problem.cpp:401-416): advect to the
// nearest face, pore wall or intrinsic event, the kill test, and the tally of
// the flight.  Returns the event (a plane index, -1 intrinsic, -3 the pore)
// and sets fate to MCG_ESCAPED when the bundle is lost; dmin and apos receive
// the distance and the position after the flight.  `walking` is set when the
// flight's 2-D / 3-D tally walk was deferred to W (FP32 flight only).
template <typename F, int kDefer>
__device__ __forceinline__ int flight_step(const KParams<F>& K, const View<F>& V, const Sink& sink,
                                           Lane<F>& L, WalkOf<F, kDefer>& W, bool& walking, uint32_t& fate,
                                           F& dmin_out, dev::V3<F>& apos, int32_t& pbeg_out,
                                           int32_t& pore_out) {
    walking = false;
    const DevSdom<F>& S = V.sdoms[L.sdom];
    const F vel = V.vel[L.mode];
    // the cell's fields the flight tests, read once and together (a lone
    // lane would otherwise wait for each read behind its own branch); the
    // collision gets the face base and the pore from here rather than read
    // them again behind the tally's LDS atomics
    const int32_t s_box = S.box, s_pore = S.pore, s_pbeg = S.pbeg;
    const F s_eps = S.eps;
    pbeg_out = s_pbeg;
    pore_out = s_pore;
    // advect: nearest plane the bundle moves towards (subdomain.cpp:161-192)
    F dmin = L.scat_next;
    int hit = -1;
    // FP32 keeps each face's signed distance and approach rate for the kill
    // test below (at most kMaxFaces faces per cell, unrolled into registers).
    // A bundle that rounding left just outside a face it moves towards meets
    // that face where it stands (parameter 0) instead of a parameter below
    // zero, which at a grazing angle would exceed the kill tolerance: the
    // face's rule (reflection, Inter crossing, absorption) applies instead of
    // a loss.  FP64 keeps the checker's arithmetic.
    F dist[kMaxFaces], rate[kMaxFaces];
    F klo[3], khi[3];   // FP32 box: the kill test's bounds, read with the faces'
    if constexpr (sizeof(F) == 4) {
        if (s_box) {
            // an axis-aligned box: on each axis only the face ahead can be
            // met, at (bound - p_a) / u_a.  The bounds and face numbers are
            // read together first and the nearest face is kept with selects,
            // so that the flight waits for one round of LDS reads rather
            // than two per axis
            const F p[3] = {L.pos.x, L.pos.y, L.pos.z}, u[3] = {L.dir.x, L.dir.y, L.dir.z};
            F lo[3], hi[3];
            int32_t flo[3], fhi[3];
            const int32_t pbeg = s_pbeg;
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                lo[a] = S.bmin[a];
                hi[a] = S.bmax[a];
                flo[a] = pbeg + S.fmin[a];
                fhi[a] = pbeg + S.fmax[a];
                klo[a] = S.kmin[a];
                khi[a] = S.kmax[a];
            }
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                const bool up = u[a] > F(0);
                const F ta = ((up ? hi[a] : lo[a]) - p[a]) * __builtin_amdgcn_rcpf(u[a]);
                const F t = ta < F(0) ? F(0) : ta;
                const bool take = u[a] != F(0) && t < dmin;
                dmin = take ? t : dmin;
                hit = take ? (up ? fhi[a] : flo[a]) : hit;
            }
        } else {
            // every lane evaluates every face: a reciprocal is cheaper than
            // the divergence of skipping the faces it moves away from.  A
            // slot past the cell's last face (every cell has at least four)
            // reads the cell's first face and is then ignored, so that the
            // face reads issue together rather than behind a branch each.
            const int32_t pbeg = s_pbeg, pend = S.pend;
#pragma unroll
            for (int i = 0; i < kMaxFaces; ++i) {
                const int b = pbeg + i;
                const bool real = b < pend;
                const DevPlane<F>& P = V.planes[real ? b : pbeg];
                const F nd = dev::dot3(P.n, L.dir);
                const F d = __builtin_fmaf(P.n[2], L.pos.z, __builtin_fmaf(P.n[1], L.pos.y,
                                                                            __builtin_fmaf(P.n[0], L.pos.x, P.c)));
                const F tf = -d * __builtin_amdgcn_rcpf(nd);
                const F t = tf < F(0) ? F(0) : tf;
                const bool take = real && nd < F(0) && t < dmin;
                dmin = take ? t : dmin;
                hit = take ? b : hit;
                dist[i] = real ? d : F(1);
                rate[i] = real ? nd : F(0);
            }
            // faces beyond the first kMaxFaces (prisms and pyramids with
            // large bases) are tested again after the move
            for (int b = s_pbeg + kMaxFaces; b < S.pend; ++b) {
                const DevPlane<F>& P = V.planes[b];
                const F nd = dev::dot3(P.n, L.dir);
                const F d = __builtin_fmaf(P.n[2], L.pos.z, __builtin_fmaf(P.n[1], L.pos.y,
                                                                            __builtin_fmaf(P.n[0], L.pos.x, P.c)));
                const F tf = -d * __builtin_amdgcn_rcpf(nd);
                const F t = tf < F(0) ? F(0) : tf;
                if (nd < F(0) && t < dmin) {
                    dmin = t;
                    hit = b;
                }
            }
        }
    } else {
        (void)dist;
        (void)rate;
        (void)klo;
        (void)khi;
        if (s_box) {
            // the checker's face loop evaluated for the face ahead on each
            // axis only: the axis's other face has n.u >= 0 (exactly: the
            // normal's other components are zeros) and the loop skips it;
            // equal parameters keep the lower face index, as the loop in face
            // order does, and a face only displaces the intrinsic event when
            // strictly nearer
            // The three faces' numbers, then their planes, are read
            // together, and the parameters computed before they are compared
            // in axis order (an axis the bundle does not move along is
            // skipped by the compare, not by a branch).
            const F u[3] = {L.dir.x, L.dir.y, L.dir.z};
            int32_t fb[3];
#pragma unroll
            for (int a = 0; a < 3; ++a) fb[a] = s_pbeg + (u[a] > F(0) ? S.fmax[a] : S.fmin[a]);
            F tt[3];
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                const DevPlane<F>& P = V.planes[fb[a]];
                tt[a] = -(P.c + dev::dot3(P.n, L.pos)) / dev::dot3(P.n, L.dir);
            }
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                const bool take = u[a] != F(0) && (tt[a] < dmin || (tt[a] == dmin && hit >= 0 && fb[a] < hit));
                dmin = take ? tt[a] : dmin;
                hit = take ? fb[a] : hit;
            }
        } else {
            for (int b = s_pbeg; b < S.pend; ++b) {
                const DevPlane<F>& P = V.planes[b];
                const F nd = dev::dot3(P.n, L.dir);
                if (nd >= F(0)) continue;
                const F t = -(P.c + dev::dot3(P.n, L.pos)) / nd;
                if (t < dmin) {
                    dmin = t;
                    hit = b;
                }
            }
        }
    }
    if (s_pore >= 0) {
        const DevPore<F>& Q = V.pores[s_pore];
        const F dx = L.pos.x - Q.cx, dy = L.pos.y - Q.cy;
        const F bq = dx * L.dir.x + dy * L.dir.y;
        if (bq < F(0)) {
            const F aq = L.dir.x * L.dir.x + L.dir.y * L.dir.y;
            const F cq = dx * dx + dy * dy - Q.r2;
            const F disc = bq * bq - aq * cq;
            if (disc >= F(0)) {
                // a start on (or, by rounding, just inside) the wall heading
                // in meets it at once
                const F t = cq <= F(0) ? F(0) : dev::Real<F>::div(cq, dev::Real<F>::sqrt_(disc) - bq);
                if (t < dmin) {
                    dmin = t;
                    hit = -3;
                }
            }
        }
    }
    L.scat_next -= dmin;
    if (L.scat_next < dev::Real<F>::kMin) L.scat_next = F(0);
    const dev::V3<F> pre = L.pos;
    L.pos = dev::V3<F>{pre.x + dmin * L.dir.x, pre.y + dmin * L.dir.y, pre.z + dmin * L.dir.z};
    L.nstep++;
    apos = L.pos;
    dmin_out = dmin;
    // kill test (subdomain.cpp:186-190): outside any face by more than eps
    bool outside;
    if constexpr (sizeof(F) == 4) {
        if (s_box) {
            // six compares joined without short-circuits
            outside = (L.pos.x < klo[0]) | (L.pos.x > khi[0]) | (L.pos.y < klo[1]) | (L.pos.y > khi[1]) |
                      (L.pos.z < klo[2]) | (L.pos.z > khi[2]);
        } else {
            // a face's distance after the move is d + dmin * rate
            outside = false;
#pragma unroll
            for (int i = 0; i < kMaxFaces; ++i) outside |= __builtin_fmaf(dmin, rate[i], dist[i]) < -s_eps;
            for (int b = s_pbeg + kMaxFaces; b < S.pend; ++b) {
                const DevPlane<F>& P = V.planes[b];
                outside |= dev::dot3(P.n, L.pos) + P.c < -s_eps;
            }
        }
        if (s_pore >= 0) outside |= !outside_pore(V.pores[s_pore], L.pos);
    } else if (s_box) {
        // the checker's test face by face: n.p + c is p_a - bmin_a or
        // bmax_a - p_a exactly (the normal's other components are zeros)
        outside = (L.pos.x - S.bmin[0] < -s_eps) | (S.bmax[0] - L.pos.x < -s_eps) |
                  (L.pos.y - S.bmin[1] < -s_eps) | (S.bmax[1] - L.pos.y < -s_eps) |
                  (L.pos.z - S.bmin[2] < -s_eps) | (S.bmax[2] - L.pos.z < -s_eps);
        if (s_pore >= 0) outside |= !outside_pore(V.pores[s_pore], L.pos);
    } else {
        outside = !inside(V.planes, S, V.pores, L.pos);
    }
    if ((dmin < -s_eps) | outside) {
        fate = MCG_ESCAPED;
    } else {
        // tally (problem.cpp:414-416 with the kinds of problem.cpp:473-637);
        // each amount is a sign times an F value, exact in F and in FP64
        const F sg = F(L.sign);
        const F dt = sg * dev::Real<F>::div(dmin, vel);
        const F dx = sg * (L.pos.x - pre.x);
        const F dy = sg * (L.pos.y - pre.y);
        const F dz = sg * (L.pos.z - pre.z);
        const int tk = K.tally;
        const bool with_t = tk == MCG_TEMP || tk == MCG_MULTI || tk == MCG_CUMTEMP;
        const F amt[kMaxRows] = {with_t ? dt : dx, with_t ? dx : dy, with_t ? dy : dz, dz};
        const int nv = tk == MCG_MULTI ? 4 : (with_t ? 1 : 3);
        int r0 = 0;
        if (tk >= MCG_CUMTEMP) {
            const int bin = int((int64_t(L.nscat) + K.step_cum - 1) / K.step_cum);
            r0 = tk == MCG_CUMTEMP ? bin : 3 * bin;
        }
        walking = accumulate<F, kDefer>(S, sink, pre, L.pos, r0, nv, amt, W);
    }
    return hit;
}

// The collision half of a scatter-step (problem.cpp:418-433): the boundary
// scatter (boundary.cpp:283-522) or intrinsic scatter (material.cpp:226-231)
// of the event `hit` found by the flight, then the termination test.  The
// bundle's subdomain is still the flight's.  Sets fate when the bundle ends.
template <typename F>
__device__ __forceinline__ void collide_step(const KParams<F>& K, const View<F>& V, int hit, Lane<F>& L,
                                             uint32_t& fate, int32_t pbeg, int32_t pore) {
    {
        // the event's Philox blocks serve every branch below; they are drawn
        // once, outside the branches, so that a wavefront whose lanes take
        // different branches evaluates them once (FP32: block 0 alone)
        const dev::Event ev{K.k0, K.k1, L.nstep, uint32_t(L.id), uint32_t(uint64_t(L.id) >> 32)};
        const uint4 b0 = ev.block(0), b1 = low_block<F>(ev, 1);
        if constexpr (sizeof(F) == 4) {
            if (hit != -3) {
                scatter_fp32<F>(K, V, pbeg, ev, b0, b1, hit, L, fate);
            } else {
                L.nbdry++;
                pore_scatter<F>(V.pores[pore], ev, b0, b1, L);
            }
        } else {
            const bool intr = hit == -1;
            const DevPlane<F>& P = V.planes[hit >= 0 ? hit : pbeg];
            int kind = P.kind;
            if (kind == MCG_PARTIAL) kind = b1.w < P.spec_thr ? MCG_SPEC : MCG_DIFF;
            if (intr || (hit >= 0 && kind == MCG_DIFF)) {
                // drawIso (intrinsic) and drawAniso (diffuse face) with each
                // lane's own arithmetic, exactly: the same uniforms, one
                // square root for sin(theta), one sin/cos(pi q), so a
                // wavefront whose lanes take both evaluates them once
                const F r = dev::Real<F>::uniform(b0.x, b1.x) * F(2) + F(-1);
                const F ar = r < F(0) ? -r : r;
                const F sinT = dev::Real<F>::sqrt_(intr ? F(1) - r * r : ar);
                const F cosT = intr ? r : dev::Real<F>::sqrt_(F(1) - ar);
                const F q = dev::Real<F>::uniform(b0.y, b1.y) * F(2) + F(-1);
                F sq, cq;
                dev::Real<F>::sincos_pi(q, sq, cq);
                const dev::V3<F> v{sinT * cq, sinT * sq, cosT};
                L.dir = dev::normalized(intr ? v : dev::matvec(P.rot, v));   // boundary.cpp:308-312
                L.nscat++;
            }
            if (intr) {
                L.mode = dev::alias_draw(V.scat, uint32_t(K.nmodes), b0.w);
                L.scat_next = draw_path<F>(ev, b0, b1, V.mfp[L.mode]);
            } else if (hit >= 0) {
                L.nbdry++;
                switch (kind) {
                    case MCG_SPEC:   // boundary.cpp:283-287
                        L.dir = dev::normalized(dev::matvec(P.refl, L.dir));
                        break;
                    case MCG_DIFF:   // direction drawn above
                        break;
                    case MCG_ISOT:   // boundary.cpp:456-460
                        fate = MCG_ABSORBED;
                        break;
                    case MCG_PERI: { // boundary.cpp:517-522
                        const dev::V3<F> q = dev::matvec(P.prot, L.pos);
                        L.pos = dev::V3<F>{q.x + P.pt[0], q.y + P.pt[1], q.z + P.pt[2]};
                        L.dir = dev::normalized(dev::matvec(P.prot, L.dir));
                        L.sdom = P.target;
                        break;
                    }
                    default: {       // MCG_INTER, boundary.cpp:349-359
                        int target = -1;
                        if (P.icnt == 1) {
                            target = P.target;
                        } else {
                            for (int q = 0; q < P.icnt; ++q) {
                                const int cand = K.inter_list[P.ibeg + q];
                                if (inside(V.planes, V.sdoms[cand], V.pores, L.pos)) {
                                    target = cand;
                                    break;
                                }
                            }
                        }
                        if (target < 0) fate = MCG_ESCAPED;
                        else L.sdom = target;
                        break;
                    }
                }
            } else {
                L.nbdry++;
                pore_scatter<F>(V.pores[pore], ev, b0, b1, L);
            }
        }
        if (fate == 0xFFFFFFFFu) {
            if (int64_t(L.nscat) >= K.maxscat) fate = MCG_MAXSCAT;
            else if (L.nstep >= K.maxloop) fate = MCG_MAXLOOP;
        }
    }
}

// The collision with the cell's face base and pore read here.
template <typename F>
__device__ __forceinline__ void collide_step(const KParams<F>& K, const View<F>& V, int hit, Lane<F>& L,
                                             uint32_t& fate) {
    const DevSdom<F>& S = V.sdoms[L.sdom];
    collide_step<F>(K, V, hit, L, fate, S.pbeg, S.pore);
}

// One scatter-step of the bundle in L (problem.cpp:401-435): the flight and
// the collision.  Returns the bundle's fate when it ends, 0xFFFFFFFF while it
// lives.  `walking` is set when the step's 2-D / 3-D tally walk was deferred
// to W (FP32 flight only).
template <typename F, int kDefer>
__device__ __forceinline__ uint32_t scatter_step(const KParams<F>& K, const View<F>& V,
                                                 const Sink& sink, Lane<F>& L, WalkOf<F, kDefer>& W,
                                                 bool& walking) {
    uint32_t fate = 0xFFFFFFFFu;
    F dmin;
    dev::V3<F> apos;
    int32_t pbeg, pore;
    const int hit = flight_step<F, kDefer>(K, V, sink, L, W, walking, fate, dmin, apos, pbeg, pore);
    if (fate == 0xFFFFFFFFu) collide_step<F>(K, V, hit, L, fate, pbeg, pore);
    if (L.id == K.trace_id && int64_t(L.nstep) <= K.trace_cap) {
        double* r = K.trace + (int64_t(L.nstep) - 1) * MCG_TRACE_WIDTH;
        r[0] = double(L.nstep);
        r[1] = double(hit);
        r[2] = double(dmin);
        r[3] = double(apos.x);
        r[4] = double(apos.y);
        r[5] = double(apos.z);
        r[6] = double(L.dir.x);
        r[7] = double(L.dir.y);
        r[8] = double(L.dir.z);
        r[9] = double(L.scat_next);
        r[10] = double(L.mode);
        r[11] = double(L.nscat);
        r[12] = double(L.sdom);
        r[13] = fate == 0xFFFFFFFFu ? -1.0 : double(fate);
        r[14] = double(L.pos.x);
        r[15] = double(L.pos.y);
        r[16] = double(L.pos.z);
    }
    if (fate != 0xFFFFFFFFu && K.c_nscat) {
        const int64_t k = L.id - K.first;
        K.c_nscat[k] = L.nscat;
        K.c_nbdry[k] = L.nbdry;
        K.c_nstep[k] = L.nstep;
        K.c_fate[k] = fate;
    }
    return fate;
}

// Whether the wavefront claims and emits this iteration (wave-uniform).
template <int kQuorum = kEmitQuorum>
__device__ __forceinline__ bool emit_now(bool need, bool busy) {
    const uint64_t m = __ballot(need);
    if (m == 0ull) return false;
    return __popcll(m) >= kQuorum || __ballot(busy) == 0ull;
}

// The wavefront's walk phase: every walking lane advances one crossing per
// round.  The phase ends when no lane walks, or when fewer than kWalkQuorum
// do while some other lane has a step to take (the walkers resume in the
// next iteration's phase).  With no step to take it runs every walk to its
// end, so the wavefront's loop always progresses (lanes waiting to emit do
// not count: emission waits for the walkers).
// U crossings of the walking lanes (a walk that ends stops there)
template <int U, typename F, int NA>
__device__ __forceinline__ void walk_some(WalkN<F, NA>& W, bool& walking, const Sink& sink) {
    if (walking && walk_next(W, sink)) walking = false;
    if constexpr (U > 1) walk_some<U - 1>(W, walking, sink);
}

template <typename F, int NA>
__device__ __forceinline__ void walk_phase(WalkN<F, NA>& W, bool& walking, bool has_work, const Sink& sink) {
    for (;;) {
        const uint64_t m = __ballot(walking);
        if (m == 0ull) break;
        // lanes with a step to take once their walk is over
        if (popc(m) < kWalkQuorum && __ballot(!walking && has_work) != 0ull) break;
        walk_some<kWalkUnroll>(W, walking, sink);
    }
}

// Wave-level id pool: lanes whose bundle ended take the next ids of the pool,
// refilled `chunk` ids at a time from *counter by one lane.  Returns the id
// offset for lanes with `need` (all lanes call, the pool stays wave-uniform).
template <typename Counter>
__device__ __forceinline__ int64_t claim(bool need, int lane, uint64_t lt_mask, int64_t& base,
                                         int64_t& lim, Counter* counter, int chunk, int64_t count = 0,
                                         int64_t tail = 0, int64_t* seen = nullptr) {
    const uint64_t mask = __ballot(need);
    if (!mask) return 0;
    const int cnt = __popcll(mask);
    const int rank = __popcll(mask & lt_mask);
    const int64_t avail = lim - base;
    int64_t mine;
    if (avail >= cnt) {
        mine = base + rank;
        base += cnt;
    } else {
        // within `tail` ids of the end a wavefront takes no more than its
        // lanes need (at least one id per lane), so the last bundles spread
        // over the wavefronts instead of waiting in a few private pools
        const int leader = __ffsll((long long)mask) - 1;
        unsigned long long fresh = 0;
        int want = chunk;
        if (lane == leader) {
            if (tail > 0) {
                const unsigned long long cur = __hip_atomic_load(counter, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (int64_t(cur) + tail >= count) want = max(cnt - int(avail), min(chunk, kWave));
            }
            fresh = (unsigned long long)atomicAdd(counter, Counter(want));
        }
        fresh = __shfl(fresh, leader);
        want = __shfl(want, leader);
        mine = rank < avail ? base + rank : int64_t(fresh) + (rank - avail);
        base = int64_t(fresh) + (cnt - avail);
        lim = int64_t(fresh) + want;
        if (seen) *seen = int64_t(fresh) + want;
    }
    return mine;
}

// Single solve: persistent grid, every lane refilled from the solve's global
// bundle counter until the range [first, first + count) is exhausted.
template <typename F, bool kLdsTables, bool kLdsTally, int kDefer>
__global__ void __launch_bounds__(1024) trajectory_kernel(KParams<F> K) {
    extern __shared__ __align__(16) unsigned char lds[];
    const int tid = threadIdx.x;
    const View<F> V = stage<F, kLdsTables, kLdsTally>(K, lds, tid, blockDim.x);
    const Sink sink{V.direct, V.diff, K.cols};
    __syncthreads();

    const int lane = tid & (kWave - 1);
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (kWave - lane));
    int64_t pool_base = 0, pool_lim = 0;
    bool active = false, walking = false;
    Lane<F> L;
    WalkOf<F, kDefer> W;
    LaneStats st;
    if constexpr (sizeof(F) == 4) {
        // FP32: each lane also holds its next bundle, emitted ahead of time
        // (N, valid while `full`), so a lane whose bundle ends continues at
        // once.  Emission runs in rounds in which every lane with an empty
        // slot emits: a round waits for kSlotQuorum empty slots (or for
        // kEmitQuorum lanes left with nothing to run, or for a wavefront with
        // nothing in flight), so its Philox and sampling work runs on many
        // lanes at a time instead of on the few that just ended.
        Lane<F> N;
        bool full = false, drained = false;
        int64_t seen = 0;   // the claim counter when this wavefront last read it
        for (;;) {
            if (!active && !walking && full) {
                L = N;
                full = false;
                active = true;
            }
            // near the end only lanes with nothing to run fill a slot, so the
            // last bundles go to idle lanes rather than queue behind live ones
            const bool near_end = K.count - seen < K.tail;
            const bool empty = !full && !drained && (!near_end || (!active && !walking));
            const uint64_t me = __ballot(empty);
            if (me != 0ull &&
                (__popcll(me) >= kSlotQuorum || __popcll(__ballot(empty && !active && !walking)) >= kEmitQuorum ||
                 __ballot(active || walking) == 0ull)) {
                const int64_t mine = claim(empty, lane, lt_mask, pool_base, pool_lim, K.next, K.chunk, K.count,
                                           K.tail, &seen);
                if (empty) {
                    if (mine < K.count) {
                        emit<F>(K, K.flux_alias, V.mfp, K.first + mine, N);
                        full = true;
                    } else {
                        drained = true;
                    }
                }
                if (!active && !walking && full) {
                    L = N;
                    full = false;
                    active = true;
                }
            }
            if (__ballot(active || walking || full || !drained) == 0ull) break;
            if (active && !walking) {
                const uint32_t fate = scatter_step<F, kDefer>(K, V, sink, L, W, walking);
                if (fate != 0xFFFFFFFFu) {
                    st.add(fate, L.nstep);
                    active = false;
                }
            }
            if constexpr (kDefer != 0) walk_phase(W, walking, active, sink);
        }
    } else {
        bool finished = false;
        for (;;) {
            const bool need = !active && !finished && !walking;
            if (emit_now<kEmitQuorumFp64>(need, active || walking)) {
                const int64_t mine = claim(need, lane, lt_mask, pool_base, pool_lim, K.next, K.chunk);
                if (need) {
                    if (mine < K.count) {
                        emit<F>(K, K.flux_alias, V.mfp, K.first + mine, L);
                        active = true;
                    } else {
                        finished = true;
                    }
                }
            }
            if (__ballot(active || walking) == 0ull) break;
            if (active && !walking) {
                const uint32_t fate = scatter_step<F, kDefer>(K, V, sink, L, W, walking);
                if (fate != 0xFFFFFFFFu) {
                    st.add(fate, L.nstep);
                    active = false;
                }
            }
            if constexpr (kDefer != 0) walk_phase(W, walking, active, sink);
        }
    }
    st.publish(K.counters, lane);
    if (kLdsTally) {
        __syncthreads();
        flush_tally(V.direct, V.diff, K.direct, K.diff, K.rows, K.cols, K.lds_diff != 0, tid, blockDim.x);
    }
}

// A trajectory from a given start (TrajProblem::solve with pos_ set,
// problem.cpp:226-299): one lane, event 0 of bundle K.trace_id for the mode
// (scattering distribution unless given), the direction (isotropic unless
// given) and the first free path, then the solve's step loop.
template <typename F>
struct TraceStart {
    double pos[3], dir[3];
    int32_t has_dir, mode, sdom, pad;
};

template <typename F>
__global__ void __launch_bounds__(64) trace_from_kernel(KParams<F> K, TraceStart<F> s) {
    extern __shared__ __align__(16) unsigned char lds[];
    const View<F> V = stage<F, false, false>(K, lds, threadIdx.x, blockDim.x);
    const Sink sink{V.direct, V.diff, K.cols};
    __syncthreads();
    if (threadIdx.x != 0) return;
    const dev::Event ev{K.k0, K.k1, 0u, uint32_t(K.trace_id), uint32_t(uint64_t(K.trace_id) >> 32)};
    const uint4 b0 = ev.block(0), b1 = low_block<F>(ev, 1);
    Lane<F> L;
    L.mode = s.mode >= 0 ? uint32_t(s.mode) : dev::alias_draw(K.scat_alias, uint32_t(K.nmodes), b0.w);
    const dev::V3<F> d = s.has_dir ? dev::V3<F>{F(s.dir[0]), F(s.dir[1]), F(s.dir[2])} : dev::draw_iso<F>(b0, b1);
    L.dir = dev::normalized(d);
    L.pos = dev::V3<F>{F(s.pos[0]), F(s.pos[1]), F(s.pos[2])};
    L.sdom = s.sdom;
    L.sign = 1;
    L.scat_next = draw_path<F>(ev, b0, b1, V.mfp[L.mode]);
    L.nscat = L.nbdry = L.nstep = 0;
    L.id = K.trace_id;
    Walk<F> W;
    bool walking = false;
    while (scatter_step<F, 0>(K, V, sink, L, W, walking) == 0xFFFFFFFFu) {
    }
}

// Many solves in one persistent launch.  Workgroups start spread over the
// solves and then visit every solve once, in order: a solve with bundles
// left is staged in LDS (tables, faces, tally) and its bundles are claimed
// from the solve's own counter, wavefront pool by wavefront pool, exactly as
// trajectory_kernel claims them, until none are left; the workgroup then
// flushes the solve's LDS tally and moves on.  A workgroup waits for its
// slowest wavefront only when it leaves a solve.  Every solve is visited by
// every workgroup, and a visit leaves the solve with no unclaimed bundles,
// so all bundles are processed; a counter read that is stale only makes a
// visit find nothing to do.
template <typename F, bool kLdsTables, bool kLdsTally, int kDefer>
__global__ void __launch_bounds__(1024) batch_kernel(const KParams<F>* __restrict__ solves, int nsolves) {
    extern __shared__ __align__(16) unsigned char lds[];
    __shared__ int s_go, s_si;
    const int tid = threadIdx.x;
    const int lane = tid & (kWave - 1);
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (kWave - lane));
    const int first = int((int64_t(blockIdx.x) * nsolves) / int64_t(gridDim.x));
    for (int v = 0; v < nsolves; ++v) {
        if (tid == 0) {
            const int si = first + v < nsolves ? first + v : first + v - nsolves;
            s_si = si;
            s_go = __hip_atomic_load(solves[si].next, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) <
                           (unsigned long long)solves[si].count ? 1 : 0;
        }
        __syncthreads();
        const bool go = s_go != 0;
        // the solve index comes through LDS: the compiler then reads the
        // solve's parameters per lane where it uses them, rather than holding
        // all of them in scalar registers (which it cannot for a batch)
        const KParams<F>& K = solves[s_si];
        if (!go) {
            __syncthreads();   // s_go is rewritten by the next visit
            continue;
        }
        const View<F> V = stage<F, kLdsTables, kLdsTally>(K, lds, tid, blockDim.x);
        const Sink sink{V.direct, V.diff, K.cols};
        __syncthreads();
        int64_t pool_base = 0, pool_lim = 0;
        bool active = false, finished = false, walking = false;
        Lane<F> L;
        WalkOf<F, kDefer> W;
        LaneStats st;
        for (;;) {
            const bool need = !active && !finished && !walking;
            if (emit_now(need, active || walking)) {
                const int64_t mine = claim(need, lane, lt_mask, pool_base, pool_lim, K.next, K.chunk);
                if (need) {
                    if (mine < K.count) {
                        emit<F>(K, K.flux_alias, V.mfp, K.first + mine, L);
                        active = true;
                    } else {
                        finished = true;
                    }
                }
            }
            if (__ballot(active || walking) == 0ull) break;
            if (active && !walking) {
                const uint32_t fate = scatter_step<F, kDefer>(K, V, sink, L, W, walking);
                if (fate != 0xFFFFFFFFu) {
                    st.add(fate, L.nstep);
                    active = false;
                }
            }
            if constexpr (kDefer != 0) walk_phase(W, walking, active, sink);
        }
        st.publish(K.counters, lane);
        __syncthreads();
        if (kLdsTally) flush_tally(V.direct, V.diff, K.direct, K.diff, K.rows, K.cols, K.lds_diff != 0, tid, blockDim.x);
        __syncthreads();
    }
}

// ================================================================ event engine
// The event-based variant of the solve: bundle state in HBM as structure of
// arrays, advanced one scatter-step per bundle per round.  A round performs
// each live bundle's pending collision (the event its last flight found),
// then its next flight and tally, and appends the survivor to the next
// round's list under the key (class of the event just found: intrinsic,
// planar face or pore wall; box cell or other), so that a wavefront of the
// next round runs one collision branch on one advect path.  Isothermal walls
// absorb at once and lost bundles leave at once.  Emission refills the list to
// a target number of live bundles; when no bundle is left to emit and the
// list is short, a drain launch finishes the rest history-style.  Every
// bundle takes exactly the flights and collisions of the history kernel, with
// the same Philox events, so results equal it bundle for bundle (tallies up
// to FP64 summation order).
constexpr int kEvGroups = 6;
constexpr int32_t kEvHole = -1000;   // pending-event mark of a slot left empty in place
constexpr int kEvPad = 32;           // counters 256 B apart: one cache line each

template <typename F>
struct EvState {
    F *px, *py, *pz, *ux, *uy, *uz, *sn;
    uint32_t* mode;
    int32_t* sds;        // subdomain << 1 | (sign < 0)
    uint32_t *nscat, *nbdry, *nstep;
    int64_t* id;
    int32_t* hit;        // pending event (plane, -1 intrinsic, -3 pore), kEvHole for an empty slot
};

template <typename F>
struct EvParams {
    KParams<F> K;
    EvState<F> in, out;                 // kEvGroups regions of `cap` slots each
    unsigned long long* in_count;       // entries per group of the input list (kEvPad apart)
    unsigned long long* out_count;      // entries per group of the output list (kEvPad apart)
    unsigned long long* emit;           // [0] first id offset, [1] bundles to emit this round
    int64_t cap;
    int64_t live_target;
    int32_t regroup;                    // 1: survivors appended to `out` by group; 0: kept in place
};

template <typename F>
__device__ __forceinline__ void ev_load(const EvState<F>& S, int64_t s, Lane<F>& L, int& hit) {
    L.pos = dev::V3<F>{S.px[s], S.py[s], S.pz[s]};
    L.dir = dev::V3<F>{S.ux[s], S.uy[s], S.uz[s]};
    L.scat_next = S.sn[s];
    L.mode = S.mode[s];
    const int32_t sd = S.sds[s];
    L.sdom = sd >> 1;
    L.sign = (sd & 1) ? -1 : 1;
    L.nscat = S.nscat[s];
    L.nbdry = S.nbdry[s];
    L.nstep = S.nstep[s];
    L.id = S.id[s];
    hit = S.hit[s];
}

template <typename F>
__device__ __forceinline__ void ev_store(const EvState<F>& S, int64_t s, const Lane<F>& L, int hit) {
    S.px[s] = L.pos.x;
    S.py[s] = L.pos.y;
    S.pz[s] = L.pos.z;
    S.ux[s] = L.dir.x;
    S.uy[s] = L.dir.y;
    S.uz[s] = L.dir.z;
    S.sn[s] = L.scat_next;
    S.mode[s] = L.mode;
    S.sds[s] = (L.sdom << 1) | (L.sign < 0 ? 1 : 0);
    S.nscat[s] = L.nscat;
    S.nbdry[s] = L.nbdry;
    S.nstep[s] = L.nstep;
    S.id[s] = L.id;
    S.hit[s] = hit;
}

// Slot of entry e of a list whose groups hold count[0..kEvGroups) entries
// (group g occupies slots [g cap, g cap + count[g]))
__device__ __forceinline__ int64_t ev_slot(const unsigned long long (&cnt)[kEvGroups], int64_t e, int64_t cap) {
    int64_t acc = 0;
#pragma unroll
    for (int g = 0; g < kEvGroups; ++g) {
        const int64_t c = int64_t(cnt[g]);
        if (e < acc + c) return int64_t(g) * cap + (e - acc);
        acc += c;
    }
    return -1;
}

// Append the lanes with `want` to group `grp` of a list; returns the lane's
// slot (or -1).  The wavefronts of a workgroup reserve their places in LDS
// and one thread per group reserves the workgroup's in the list, so a
// workgroup costs one global atomic per group.  Every thread of the
// workgroup calls (three barriers).
struct EvAppendLds {
    unsigned int cnt[kEvGroups];
    unsigned long long base[kEvGroups];
};

__device__ __forceinline__ int64_t ev_append(unsigned long long* count, int64_t cap, int grp, bool want, int lane,
                                             uint64_t lt_mask, EvAppendLds& A) {
    const int tid = threadIdx.x;
    if (tid < kEvGroups) A.cnt[tid] = 0u;
    __syncthreads();
    unsigned int off = 0;
#pragma unroll
    for (int g = 0; g < kEvGroups; ++g) {
        const uint64_t m = __ballot(want && grp == g);
        if (m == 0ull) continue;
        const int leader = __ffsll((long long)m) - 1;
        unsigned int w = 0;
        if (lane == leader) w = atomicAdd(&A.cnt[g], unsigned(popc(m)));
        w = __shfl(w, leader);
        if (want && grp == g) off = w + unsigned(popc(m & lt_mask));
    }
    __syncthreads();
    if (tid < kEvGroups) A.base[tid] = A.cnt[tid] ? atomicAdd(count + tid * kEvPad, (unsigned long long)A.cnt[tid]) : 0ull;
    __syncthreads();
    return want ? int64_t(grp) * cap + int64_t(A.base[grp]) + int64_t(off) : -1;
}

// The next flight of a bundle whose collision left it alive: advect, kill
// test, tally (walks inline), and an isothermal wall's absorption at once.
template <typename F>
__device__ __forceinline__ int ev_flight(const KParams<F>& K, const View<F>& V, const Sink& sink, Lane<F>& L,
                                         uint32_t& fate) {
    Walk<F> W;
    bool walking = false;
    F dmin;
    dev::V3<F> apos;
    int32_t pbeg, pore;
    const int hit = flight_step<F, 0>(K, V, sink, L, W, walking, fate, dmin, apos, pbeg, pore);
    if (fate == 0xFFFFFFFFu && hit >= 0 && V.planes[hit].kind == MCG_ISOT) collide_step<F>(K, V, hit, L, fate);
    return hit;
}

template <typename F>
__device__ __forceinline__ int ev_group(const View<F>& V, const Lane<F>& L, int hit) {
    const int cls = hit == -1 ? 0 : (hit == -3 ? 2 : 1);
    return 2 * cls + (V.sdoms[L.sdom].box ? 0 : 1);
}

template <typename F>
__device__ __forceinline__ void ev_finish(const KParams<F>& K, const Lane<F>& L, uint32_t fate, LaneStats& st) {
    st.add(fate, L.nstep);
    if (K.c_nscat) {
        const int64_t k = L.id - K.first;
        K.c_nscat[k] = L.nscat;
        K.c_nbdry[k] = L.nbdry;
        K.c_nstep[k] = L.nstep;
        K.c_fate[k] = fate;
    }
}

// One round over the input list.
template <typename F, bool kLdsTally>
__global__ void __launch_bounds__(1024) ev_round_kernel(EvParams<F> E) {
    extern __shared__ __align__(16) unsigned char lds[];
    __shared__ EvAppendLds app;
    const KParams<F>& K = E.K;
    const int tid = threadIdx.x;
    const View<F> V = stage<F, false, kLdsTally>(K, lds, tid, blockDim.x);
    const Sink sink{V.direct, V.diff, K.cols};
    __syncthreads();
    const int lane = tid & (kWave - 1);
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (kWave - lane));
    unsigned long long cnt[kEvGroups];
    int64_t total = 0;
#pragma unroll
    for (int g = 0; g < kEvGroups; ++g) {
        cnt[g] = E.in_count[g * kEvPad];
        total += int64_t(cnt[g]);
    }
    LaneStats st;
    const int64_t stride = int64_t(gridDim.x) * blockDim.x;
    for (int64_t base = int64_t(blockIdx.x) * blockDim.x; base < total; base += stride) {
        const int64_t e = base + tid;
        const int64_t s = e < total ? ev_slot(cnt, e, E.cap) : -1;
        Lane<F> L;
        int hit = kEvHole;
        if (s >= 0) ev_load(E.in, s, L, hit);
        bool live = hit != kEvHole;
        int grp = 0;
        if (live) {
            uint32_t fate = 0xFFFFFFFFu;
            collide_step<F>(K, V, hit, L, fate);
            if (fate == 0xFFFFFFFFu) hit = ev_flight<F>(K, V, sink, L, fate);
            if (fate != 0xFFFFFFFFu) {
                ev_finish(K, L, fate, st);
                live = false;
            } else {
                grp = ev_group(V, L, hit);
            }
        }
        if (E.regroup) {
            const int64_t o = ev_append(E.out_count, E.cap, grp, live, lane, lt_mask, app);
            if (live) ev_store(E.out, o, L, hit);
        } else if (s >= 0) {
            // kept in place: a bundle that ended leaves a hole
            if (live) ev_store(E.in, s, L, hit);
            else if (hit != kEvHole) E.in.hit[s] = kEvHole;
        }
    }
    st.publish(K.counters, lane);
    if (kLdsTally) {
        __syncthreads();
        flush_tally(V.direct, V.diff, K.direct, K.diff, K.rows, K.cols, K.lds_diff != 0, tid, blockDim.x);
    }
}

// How many bundles this round emits (up to the live target, within the
// bundles left), and the clearing of the counters the next round appends to.
__global__ void ev_prep_kernel(unsigned long long* out_count, unsigned long long* next_count,
                               unsigned long long* next_id, unsigned long long* emit, int64_t count,
                               int64_t live_target) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    int64_t live = 0;
    for (int g = 0; g < kEvGroups; ++g) live += int64_t(out_count[g * kEvPad]);
    const int64_t left = count - int64_t(*next_id);
    const int64_t want = live_target > live ? live_target - live : 0;
    const int64_t n = want < left ? want : left;
    emit[0] = *next_id;
    emit[1] = (unsigned long long)(n > 0 ? n : 0);
    *next_id += emit[1];
    for (int g = 0; g < kEvGroups; ++g) next_count[g * kEvPad] = 0ull;
}

// Emission of this round's new bundles and their first flight; appended to
// the output list.
template <typename F, bool kLdsTally>
__global__ void __launch_bounds__(1024) ev_emit_kernel(EvParams<F> E) {
    extern __shared__ __align__(16) unsigned char lds[];
    __shared__ EvAppendLds app;
    const KParams<F>& K = E.K;
    const int tid = threadIdx.x;
    const View<F> V = stage<F, false, kLdsTally>(K, lds, tid, blockDim.x);
    const Sink sink{V.direct, V.diff, K.cols};
    __syncthreads();
    const int lane = tid & (kWave - 1);
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (kWave - lane));
    const int64_t first = int64_t(E.emit[0]), n = int64_t(E.emit[1]);
    LaneStats st;
    const int64_t stride = int64_t(gridDim.x) * blockDim.x;
    for (int64_t base = int64_t(blockIdx.x) * blockDim.x; base < n; base += stride) {
        const int64_t t = base + tid;
        Lane<F> L;
        bool live = false;
        int hit = 0, grp = 0;
        if (t < n) {
            emit<F>(K, K.flux_alias, V.mfp, K.first + first + t, L);
            uint32_t fate = 0xFFFFFFFFu;
            hit = ev_flight<F>(K, V, sink, L, fate);
            if (fate != 0xFFFFFFFFu) {
                ev_finish(K, L, fate, st);
            } else {
                live = true;
                grp = ev_group(V, L, hit);
            }
        }
        const int64_t o = ev_append(E.out_count, E.cap, grp, live, lane, lt_mask, app);
        if (live) ev_store(E.out, o, L, hit);
    }
    st.publish(K.counters, lane);
    if (kLdsTally) {
        __syncthreads();
        flush_tally(V.direct, V.diff, K.direct, K.diff, K.rows, K.cols, K.lds_diff != 0, tid, blockDim.x);
    }
}

// The tail: every bundle left in the list runs to its end, history-style.
template <typename F, bool kLdsTally>
__global__ void __launch_bounds__(1024) ev_drain_kernel(EvParams<F> E) {
    extern __shared__ __align__(16) unsigned char lds[];
    const KParams<F>& K = E.K;
    const int tid = threadIdx.x;
    const View<F> V = stage<F, false, kLdsTally>(K, lds, tid, blockDim.x);
    const Sink sink{V.direct, V.diff, K.cols};
    __syncthreads();
    const int lane = tid & (kWave - 1);
    unsigned long long cnt[kEvGroups];
    int64_t total = 0;
#pragma unroll
    for (int g = 0; g < kEvGroups; ++g) {
        cnt[g] = E.in_count[g * kEvPad];
        total += int64_t(cnt[g]);
    }
    LaneStats st;
    const int64_t stride = int64_t(gridDim.x) * blockDim.x;
    for (int64_t e = int64_t(blockIdx.x) * blockDim.x + tid; e < total; e += stride) {
        const int64_t s = ev_slot(cnt, e, E.cap);
        Lane<F> L;
        int hit = kEvHole;
        ev_load(E.in, s, L, hit);
        if (hit == kEvHole) continue;
        uint32_t fate = 0xFFFFFFFFu;
        for (;;) {
            collide_step<F>(K, V, hit, L, fate);
            if (fate != 0xFFFFFFFFu) break;
            hit = ev_flight<F>(K, V, sink, L, fate);
            if (fate != 0xFFFFFFFFu) break;
        }
        ev_finish(K, L, fate, st);
    }
    st.publish(K.counters, lane);
    if (kLdsTally) {
        __syncthreads();
        flush_tally(V.direct, V.diff, K.direct, K.diff, K.rows, K.cols, K.lds_diff != 0, tid, blockDim.x);
    }
}

// ------------------------------------------------------------------ host
enum class KernelKind { Trajectory, Batch };

template <typename F, bool kTables, bool kTally, int kDefer>
const void* kernel_of(KernelKind kind) {
    // FP64 defers walks in single solves only: no such batch kernel is built
    if constexpr (sizeof(F) == 8 && kDefer != 0)
        return reinterpret_cast<const void*>(trajectory_kernel<F, kTables, kTally, kDefer>);
    else
        return kind == KernelKind::Trajectory
                   ? reinterpret_cast<const void*>(trajectory_kernel<F, kTables, kTally, kDefer>)
                   : reinterpret_cast<const void*>(batch_kernel<F, kTables, kTally, kDefer>);
}

template <typename F, int kDefer>
const void* kernel_by_plan(KernelKind kind, bool tables, bool tally) {
    return tables ? (tally ? kernel_of<F, true, true, kDefer>(kind) : kernel_of<F, true, false, kDefer>(kind))
                  : (tally ? kernel_of<F, false, true, kDefer>(kind) : kernel_of<F, false, false, kDefer>(kind));
}

// Grids smaller than this walk inside the step unless deferral is asked for:
// a walk over a few cells (the octet truss's struts, 4-8 cells) ends in a
// crossing or two, and the walk phase's set-up and checks cost more than the
// idle lanes they save (FP32 octet 35.4 ms in-step, 37.8 ms deferred; a
// 7 x 5 x 3 map of the porous film 36.5 ms in-step, 34.8 ms deferred).
constexpr int64_t kDeferMinCells = 32;

// Axes of a launch's deferred tally walks (0: each walk runs inside its
// step).  Only scenes with 2-D / 3-D grids defer, over two axes when no grid
// is 3-D; FP64 only in single solves (no FP64 batch kernel that defers is
// built).  `knob` is mcg_tune.defer_walks: -1 never, 1 wherever such a
// kernel exists, 0 when some grid has kDeferMinCells cells.
template <typename F>
int walk_axes(bool grid, bool grid_2d, int64_t cells, KernelKind kind, int knob) {
    if (!grid || knob < 0 || (knob == 0 && cells < kDeferMinCells)) return 0;
    if (sizeof(F) == 8 && kind != KernelKind::Trajectory) return 0;
    return grid_2d ? 2 : 3;
}

// The kernel variant for an LDS plan and walk_axes' choice.
template <typename F>
const void* pick_kernel(KernelKind kind, bool tables, bool tally, int walk) {
    if constexpr (sizeof(F) == 4) {
        if (walk == 2) return kernel_by_plan<F, 2>(kind, tables, tally);
        if (walk == 3) return kernel_by_plan<F, 3>(kind, tables, tally);
    } else {
        if (kind == KernelKind::Trajectory && walk == 2) return kernel_by_plan<F, 2>(kind, tables, tally);
        if (kind == KernelKind::Trajectory && walk == 3) return kernel_by_plan<F, 3>(kind, tables, tally);
    }
    return kernel_by_plan<F, 0>(kind, tables, tally);
}

thread_local std::string t_err;

#define HIP_TRY(expr)                                                                   \
    do {                                                                                \
        hipError_t e_ = (expr);                                                         \
        if (e_ != hipSuccess)                                                           \
            return ::mcg::set_error(MCG_EDEVICE, std::string(#expr ": ") + hipGetErrorString(e_)); \
    } while (0)

template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    bool own = true;   // false: a view into memory owned elsewhere
    ~DevBuf() { if (p && own) (void)hipFree(p); }
    // `count` elements at `ptr`, owned by someone else
    void view(T* ptr, size_t count) {
        if (p && own) (void)hipFree(p);
        p = ptr;
        n = count;
        own = false;
    }
    int alloc(size_t count) {
        if (p && own) (void)hipFree(p);
        p = nullptr;
        own = true;
        n = count;
        if (count == 0) return MCG_OK;
        hipError_t e = hipMalloc(&p, count * sizeof(T));
        if (e != hipSuccess) { p = nullptr; return set_error(MCG_ENOMEM, std::string("hipMalloc: ") + hipGetErrorString(e)); }
        return MCG_OK;
    }
    // at least `count` elements, reusing the allocation when it is large enough
    int reserve(size_t count) {
        if (p && count <= n) return MCG_OK;
        return alloc(count);
    }
    int upload(const T* src, size_t count) {
        int rc = alloc(count);
        if (rc || count == 0) return rc;
        hipError_t e = hipMemcpy(p, src, count * sizeof(T), hipMemcpyHostToDevice);
        if (e != hipSuccess) return set_error(MCG_EDEVICE, std::string("hipMemcpy: ") + hipGetErrorString(e));
        return MCG_OK;
    }
};

template <typename F>
struct DevScene {
    DevBuf<DevPlane<F>> planes;
    DevBuf<DevSdom<F>> sdoms;
    DevBuf<DevEmitter<F>> emitters;
    DevBuf<DevPore<F>> pores;
    DevBuf<F> mode_vel, mode_mfp;
};

template <typename F, typename S>
void convert(const S& src, F& dst);

template <typename F>
void convert_plane(const DevPlane<double>& s, DevPlane<F>& d) {
    for (int i = 0; i < 3; ++i) { d.n[i] = F(s.n[i]); d.pt[i] = F(s.pt[i]); }
    d.c = F(s.c);
    for (int i = 0; i < 9; ++i) { d.rot[i] = F(s.rot[i]); d.refl[i] = F(s.refl[i]); d.prot[i] = F(s.prot[i]); }
    d.kind = s.kind; d.sdom = s.sdom; d.target = s.target; d.ibeg = s.ibeg; d.icnt = s.icnt; d.spec_thr = s.spec_thr;
}

template <typename F>
void convert_sdom(const DevSdom<double>& s, DevSdom<F>& d, double eps_scale) {
    for (int i = 0; i < 3; ++i) { d.o[i] = F(s.o[i]); d.divf[i] = F(s.divf[i]); d.max[i] = s.max[i]; d.stride[i] = s.stride[i]; }
    for (int i = 0; i < 9; ++i) d.inv[i] = F(s.inv[i]);
    d.eps = F(s.eps * eps_scale);
    d.accum = s.accum; d.pbeg = s.pbeg; d.pend = s.pend; d.col_off = s.col_off; d.pore = s.pore;
    d.box = s.box;
    d.gax[0] = s.gax[0];
    d.gax[1] = s.gax[1];
    for (int i = 0; i < 3; ++i) {
        d.bmin[i] = F(s.box ? s.bmin[i] : 0.0);
        d.bmax[i] = F(s.box ? s.bmax[i] : 0.0);
        d.kmin[i] = s.box ? d.bmin[i] - d.eps : F(0);
        d.kmax[i] = s.box ? d.bmax[i] + d.eps : F(0);
        d.fmin[i] = s.box ? s.fmin[i] : 0;
        d.fmax[i] = s.box ? s.fmax[i] : 0;
    }
}

template <typename F>
void convert_emitter(const DevEmitter<double>& s, DevEmitter<F>& d) {
    for (int i = 0; i < 3; ++i) { d.o[i] = F(s.o[i]); d.si[i] = F(s.si[i]); d.sj[i] = F(s.sj[i]); d.grad[i] = F(s.grad[i]); }
    for (int i = 0; i < 9; ++i) { d.mat[i] = F(s.mat[i]); d.rot[i] = F(s.rot[i]); }
    d.volume = s.volume; d.shape = s.shape; d.sdom = s.sdom; d.sign = s.sign;
    d.npiece = s.npiece;
    for (int k = 0; k < kMaxPieces; ++k) {
        d.pcum[k] = F(s.pcum[k]);
        for (int r = 0; r < 3; ++r) { d.pa[k][r] = F(s.pa[k][r]); d.pb[k][r] = F(s.pb[k][r]); }
    }
}

template <typename F>
void convert_pore(const DevPore<double>& s, DevPore<F>& d, double eps_scale) {
    d.cx = F(s.cx); d.cy = F(s.cy); d.r = F(s.r); d.r2 = F(s.r2); d.eps = F(s.eps * eps_scale);
    d.kind = s.kind; d.spec_thr = s.spec_thr;
}

template <typename F>
int upload_scene(const HostScene& h, DevScene<F>& d) {
    // The FP32 flight path scales the kill tolerance with FLT_EPSILON
    // instead of DBL_EPSILON (100 eps |M|, subdomain.cpp:45).
    const double eps_scale = sizeof(F) == 4 ? double(FLT_EPSILON) / DBL_EPSILON : 1.0;
    std::vector<DevPlane<F>> pl(h.planes.size());
    for (size_t i = 0; i < pl.size(); ++i) convert_plane(h.planes[i], pl[i]);
    std::vector<DevSdom<F>> sd(h.sdoms.size());
    for (size_t i = 0; i < sd.size(); ++i) convert_sdom(h.sdoms[i], sd[i], eps_scale);
    std::vector<DevEmitter<F>> em(h.emitters.size());
    for (size_t i = 0; i < em.size(); ++i) convert_emitter(h.emitters[i], em[i]);
    std::vector<DevPore<F>> po(h.pores.size());
    for (size_t i = 0; i < po.size(); ++i) convert_pore(h.pores[i], po[i], eps_scale);
    std::vector<F> mv(h.mode_vel.size()), mm(h.mode_mfp.size());
    for (size_t i = 0; i < mv.size(); ++i) { mv[i] = F(h.mode_vel[i]); mm[i] = F(h.mode_mfp[i]); }
    int rc;
    if ((rc = d.planes.upload(pl.data(), pl.size()))) return rc;
    if ((rc = d.sdoms.upload(sd.data(), sd.size()))) return rc;
    if ((rc = d.emitters.upload(em.data(), em.size()))) return rc;
    if ((rc = d.pores.upload(po.data(), po.size()))) return rc;
    if ((rc = d.mode_vel.upload(mv.data(), mv.size()))) return rc;
    if ((rc = d.mode_mfp.upload(mm.data(), mm.size()))) return rc;
    return MCG_OK;
}

}  // namespace

int set_error(int code, const std::string& msg) {
    t_err = msg;
    return code;
}

}  // namespace mcg

struct mcg_scene {
    mcg::HostScene host;
    mcg::DevScene<double> d64;
    mcg::DevScene<float> d32;
    mcg::DevBuf<uint64_t> scat_alias, flux_alias;
    mcg::DevBuf<int32_t> inter_list;
    int device = 0;
    int num_cu = 0;
    size_t lds_limit = 0;
    mcg_tune tune{};                // launch shape and engine of this scene's solves (0 = default)
    // per-solve device workspace, kept between solves (one thread solves a
    // scene at a time): tallies (direct, difference array, then the
    // counters, read back in one copy into pinned host memory), emitter CDF
    // (uploaded again only when it changes), timing events
    mcg::DevBuf<double> ws_tally;
    mcg::DevBuf<double> ws_priv;     // private copies of a tally left in global memory (one per XCD but the first)
    mcg::DevBuf<int64_t> ws_cdf;
    std::vector<int64_t> ws_cdf_host;
    double* pinned = nullptr;
    size_t pinned_n = 0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // event engine: two bundle lists (structure of arrays) and their counters
    mcg::DevBuf<unsigned char> ev_state;
    mcg::DevBuf<unsigned long long> ev_ctl;
    ~mcg_scene() {
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (pinned) (void)hipHostFree(pinned);
    }
};

namespace mcg {
namespace {

template <typename F>
size_t lds_bytes(const HostScene& h, int rows, bool tables, bool tally) {
    size_t off = ((tables ? h.mode_cell.size() : 0) * (8 + 2 * sizeof(F)) + 15) & ~size_t(15);
    off = (off + sizeof(DevPlane<F>) * h.planes.size() + 15) & ~size_t(15);
    off = (off + sizeof(DevSdom<F>) * h.sdoms.size() + 15) & ~size_t(15);
    off = (off + sizeof(DevPore<F>) * h.pores.size() + 15) & ~size_t(15);
    if (tally) off += sizeof(double) * (size_t(rows) * h.cols + (h.has_1d ? size_t(rows) * (h.cols + 1) : 0));
    return off;
}

// What a launch stages in LDS.  The tables and the tally both go in when
// they fit; a 1-D tally only while it is small (it takes O(1) adds per step
// through range adds, so LDS spent on it costs occupancy for little), a 2-D
// / 3-D map whatever its size (a step adds to every cell it crosses, and
// global FP64 atomics for those bound the whole solve), at the price of the
// tables, which then stay in global memory behind the L2.
struct LdsPlan {
    bool tables = false, tally = false;
    size_t bytes = 0;
};

LdsPlan plan_lds(size_t geo, size_t tab, size_t tally, bool grid, bool tables_allowed, size_t limit) {
    const bool tab_fits = tables_allowed && geo + tab <= limit;
    const bool tally_ok = grid || tally <= kLdsTallyCap;
    if (tab_fits && tally_ok && geo + tab + tally <= limit) return {true, true, geo + tab + tally};
    if (grid && geo + tally <= limit) return {false, true, geo + tally};
    if (tab_fits) return {true, false, geo + tab};
    if (tally_ok && geo + tally <= limit) return {false, true, geo + tally};
    return {false, false, geo};
}

template <typename F>
DevScene<F>& dev_scene(mcg_scene* sc) {
    return sizeof(F) == 8 ? reinterpret_cast<DevScene<F>&>(sc->d64)
                          : reinterpret_cast<DevScene<F>&>(sc->d32);
}

// Scene and problem fields of a solve's kernel parameters (buffers aside).
template <typename F>
void fill_params(mcg_scene* sc, const HostProblem& p, KParams<F>& K) {
    const HostScene& h = sc->host;
    DevScene<F>& d = dev_scene<F>(sc);
    K.planes = d.planes.p;
    K.sdoms = d.sdoms.p;
    K.emitters = d.emitters.p;
    K.pores = d.pores.p;
    K.inter_list = sc->inter_list.p;
    K.mode_vel = d.mode_vel.p;
    K.mode_mfp = d.mode_mfp.p;
    K.scat_alias = sc->scat_alias.p;
    K.flux_alias = sc->flux_alias.p;
    K.nplanes = int32_t(h.planes.size());
    K.nsdom = int32_t(h.sdoms.size());
    K.nemitter = int32_t(h.emitters.size());
    K.npore = int32_t(h.pores.size());
    K.nmodes = int32_t(h.mode_cell.size());
    K.tally = p.tally;
    K.rows = int32_t(p.rows);
    K.cols = int32_t(h.cols);
    K.step_cum = p.step > 0 ? p.step : 1;
    K.maxscat = p.maxscat;
    K.maxloop = uint32_t(std::min<int64_t>(p.maxloop, 0xFFFFFFFEll));
    K.k0 = uint32_t(p.seed);
    K.k1 = uint32_t(p.seed >> 32);
    K.first = p.first;
    K.count = p.count;
    K.chunk = sc->tune.chunk > 0 ? sc->tune.chunk : kChunk;
    K.lds_diff = h.has_1d ? 1 : 0;
    K.trace_id = -1;
}

// Clears tallies and counters: one launch instead of several memsets, and
// usable inside a captured graph (memset nodes in captured graphs are
// avoided).
__global__ void zero_kernel(double* a, size_t na, double* b, size_t nb, unsigned long long* c, size_t nc) {
    const size_t stride = size_t(gridDim.x) * blockDim.x;
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < na + nb + nc; i += stride) {
        if (i < na) a[i] = 0.0;
        else if (i < na + nb) b[i - na] = 0.0;
        else c[i - na - nb] = 0ull;
    }
}

// Adds the private copies of a global tally into the first: a[i] +=
// sum over c of priv[c * stride + i], i < n.
__global__ void fold_copies_kernel(double* a, size_t n, const double* priv, size_t stride, int ncopies) {
    const size_t step = size_t(gridDim.x) * blockDim.x;
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += step) {
        double s = 0.0;
        for (int c = 0; c < ncopies; ++c) s += priv[size_t(c) * stride + i];
        a[i] += s;
    }
}

// Private copies of a tally that stays in global memory (one per XCD)
constexpr int kTallyCopies = 8;

// Launch shape and LDS plan of a single solve, then the launch on `stream`
// (K holds the solve's buffers; its LDS and tail fields are set here).
template <typename F>
int launch_solve(mcg_scene* sc, const HostProblem& p, KParams<F>& K, hipStream_t stream,
                 bool may_allocate = false) {
    const HostScene& h = sc->host;
    const int rows = int(p.rows);
    const size_t geo = lds_bytes<F>(h, rows, false, false);
    const LdsPlan lp = plan_lds(geo, lds_bytes<F>(h, rows, true, false) - geo,
                                lds_bytes<F>(h, rows, false, true) - geo, h.has_grid,
                                sc->tune.lds_tables >= 0, sc->lds_limit);
    const bool tables = lp.tables, tally = lp.tally;
    const size_t lds = lp.bytes;
    if (lds > sc->lds_limit) return set_error(MCG_ERANGE, "geometry exceeds the LDS budget");
    K.lds_tally = tally ? 1 : 0;
    const void* fn = pick_kernel<F>(KernelKind::Trajectory, tables, tally,
                                    walk_axes<F>(h.has_grid, h.grid_2d, h.grid_cells, KernelKind::Trajectory,
                                                 sc->tune.defer_walks));
    HIP_TRY(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)));
    int blocks_per_cu = std::max<int>(1, int(sc->lds_limit / std::max<size_t>(lds, 1)));
    // one workgroup per CU (a large LDS tally): give it the CU's 16 waves
    const int block = sc->tune.block > 0 ? sc->tune.block : (blocks_per_cu == 1 ? 1024 : 512);
    blocks_per_cu = std::min(blocks_per_cu, std::max(1, 2048 / block));
    if (sc->tune.blocks_per_cu > 0) blocks_per_cu = std::min(blocks_per_cu, int(sc->tune.blocks_per_cu));
    const int64_t needed = (p.count + block - 1) / block;
    const int grid = int(std::max<int64_t>(1, std::min<int64_t>(int64_t(sc->num_cu) * blocks_per_cu, needed)));
    K.tail = int64_t(MCG_TAIL_LANES) * grid * block;
    // a tally left in global memory gets a private copy per XCD (a solve
    // that may allocate, whose direct and diff arrays are contiguous),
    // summed into the first after the launch
    const size_t nd = size_t(p.rows) * size_t(h.cols), nf = size_t(p.rows) * size_t(h.cols + 1);
    const int ncopies = std::min(kTallyCopies, grid);
    K.ncopies = 1;
    K.priv = nullptr;
    K.copy_stride = 0;
    if (!tally && may_allocate && ncopies > 1 && K.diff == K.direct + nd) {
        int rc;
        if ((rc = sc->ws_priv.reserve(size_t(ncopies - 1) * (nd + nf)))) return rc;
        const size_t nz = size_t(ncopies - 1) * (nd + nf);
        hipLaunchKernelGGL(zero_kernel, dim3(int(std::min<size_t>(1024, (nz + 255) / 256 + 1))), dim3(256), 0, stream,
                           sc->ws_priv.p, nz, static_cast<double*>(nullptr), size_t(0),
                           static_cast<unsigned long long*>(nullptr), size_t(0));
        HIP_TRY(hipGetLastError());
        K.ncopies = ncopies;
        K.priv = sc->ws_priv.p;
        K.copy_stride = int64_t(nd + nf);
    }
    void* args[] = {&K};
    HIP_TRY(hipLaunchKernel(fn, dim3(grid), dim3(block), args, lds, stream));
    HIP_TRY(hipGetLastError());
    if (K.ncopies > 1) {
        hipLaunchKernelGGL(fold_copies_kernel, dim3(int(std::min<size_t>(1024, (nd + nf + 255) / 256))), dim3(256), 0,
                           stream, K.direct, nd + nf, static_cast<const double*>(K.priv), nd + nf, K.ncopies - 1);
        HIP_TRY(hipGetLastError());
    }
    return MCG_OK;
}

template <typename F>
int run_solve(mcg_scene* sc, const HostProblem& p, double* out, mcg_counts* counts, mcg_stats* stats,
              int64_t trace_id = -1, double* trace = nullptr, int64_t trace_cap = 0) {
    const HostScene& h = sc->host;
    auto t0 = std::chrono::steady_clock::now();
    HIP_TRY(hipSetDevice(sc->device));
    const int rows = int(p.rows);
    const int cols = int(h.cols);
    if (p.rows > (1 << 20) || h.cols * p.rows > (1ll << 31)) return set_error(MCG_ERANGE, "tally too large");
    if (p.maxloop <= 0) return set_error(MCG_EINVAL, "maxloop must be positive");
    const size_t nd = size_t(rows) * cols, nf = size_t(rows) * (cols + 1);
    DevBuf<uint32_t> cnt;
    int rc;
    const size_t nall = nd + nf + C_COUNT;   // counters: C_COUNT 8-byte words after the tallies
    if ((rc = sc->ws_tally.reserve(nall))) return rc;
    if (sc->ws_cdf.p == nullptr || sc->ws_cdf.n < p.emit_cdf.size() || sc->ws_cdf_host != p.emit_cdf) {
        if ((rc = sc->ws_cdf.reserve(p.emit_cdf.size()))) return rc;
        HIP_TRY(hipMemcpy(sc->ws_cdf.p, p.emit_cdf.data(), p.emit_cdf.size() * sizeof(int64_t),
                          hipMemcpyHostToDevice));
        sc->ws_cdf_host = p.emit_cdf;
    }
    if (sc->pinned_n < nall) {
        if (sc->pinned) HIP_TRY(hipHostFree(sc->pinned));
        sc->pinned = nullptr;
        sc->pinned_n = 0;
        HIP_TRY(hipHostMalloc(reinterpret_cast<void**>(&sc->pinned), nall * sizeof(double), hipHostMallocDefault));
        sc->pinned_n = nall;
    }
    double* const direct = sc->ws_tally.p;
    double* const diff = direct + nd;
    unsigned long long* const ctr = reinterpret_cast<unsigned long long*>(direct + nd + nf);
    const bool want_counts = counts && counts->nscat && counts->nbdry && counts->nstep && counts->fate;
    if (want_counts && (rc = cnt.alloc(size_t(4) * p.count))) return rc;
    hipLaunchKernelGGL(zero_kernel, dim3(int(std::min<size_t>(1024, (nd + nf + 255) / 256 + 1))), dim3(256), 0, 0,
                       direct, nd + nf, static_cast<double*>(nullptr), size_t(0), ctr, size_t(C_COUNT));
    HIP_TRY(hipGetLastError());

    KParams<F> K{};
    fill_params<F>(sc, p, K);
    K.emit_cdf = sc->ws_cdf.p;
    K.direct = direct;
    K.diff = diff;
    K.counters = ctr;
    K.next = ctr + C_NEXT;
    if (want_counts) {
        K.c_nscat = cnt.p;
        K.c_nbdry = cnt.p + p.count;
        K.c_nstep = cnt.p + 2 * p.count;
        K.c_fate = cnt.p + 3 * p.count;
    }
    DevBuf<double> tbuf;
    if (trace && trace_cap > 0) {
        if ((rc = tbuf.alloc(size_t(trace_cap) * MCG_TRACE_WIDTH))) return rc;
        HIP_TRY(hipMemset(tbuf.p, 0, tbuf.n * sizeof(double)));
        K.trace_id = trace_id;
        K.trace = tbuf.p;
        K.trace_cap = trace_cap;
    }
    if (!sc->ev0) HIP_TRY(hipEventCreate(&sc->ev0));
    if (!sc->ev1) HIP_TRY(hipEventCreate(&sc->ev1));
    HIP_TRY(hipEventRecord(sc->ev0, 0));
    if ((rc = launch_solve<F>(sc, p, K, 0, true))) return rc;
    HIP_TRY(hipEventRecord(sc->ev1, 0));
    HIP_TRY(hipMemcpyAsync(sc->pinned, direct, nall * sizeof(double), hipMemcpyDeviceToHost, 0));
    HIP_TRY(hipStreamSynchronize(0));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, sc->ev0, sc->ev1));

    const double* hd = sc->pinned;
    const double* hf = sc->pinned + nd;
    unsigned long long hc[C_COUNT];
    std::memcpy(hc, sc->pinned + nd + nf, sizeof(hc));
    if (want_counts) {
        HIP_TRY(hipMemcpy(counts->nscat, cnt.p, p.count * 4, hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(counts->nbdry, cnt.p + p.count, p.count * 4, hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(counts->nstep, cnt.p + 2 * p.count, p.count * 4, hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(counts->fate, cnt.p + 3 * p.count, p.count * 4, hipMemcpyDeviceToHost));
    }
    if (trace && trace_cap > 0)
        HIP_TRY(hipMemcpy(trace, tbuf.p, tbuf.n * sizeof(double), hipMemcpyDeviceToHost));
    finalise_field(h, p, hd, hf, out);
    if (stats) {
        stats->bundles = int64_t(hc[C_BUNDLES]);
        stats->steps = int64_t(hc[C_STEPS]);
        stats->escapes = int64_t(hc[C_ESC]);
        stats->absorbed = int64_t(hc[C_ABS]);
        stats->truncated = int64_t(hc[C_TRUNC]);
        stats->kernel_ms = ms;
        stats->wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    if (int64_t(hc[C_BUNDLES]) != p.count)
        return set_error(MCG_EDEVICE, "kernel finished with bundles unaccounted for");
    return MCG_OK;
}

// ------------------------------------------------- the solve that allocates nothing
// Resolves the range adds into the solve's output layout: sol (rows x cols,
// column-major) = direct + the running sum of the difference array along
// each row (both row-major).  One wavefront per row, the row scanned in
// tiles of 64 columns.
__global__ void __launch_bounds__(64) fold_kernel(const double* direct, const double* diff, int rows, int cols,
                                                  double* sol) {
    const int r = blockIdx.x;
    const int lane = threadIdx.x;
    double carry = 0.0;
    for (int c0 = 0; c0 < cols; c0 += kWave) {
        const int c = c0 + lane;
        double v = c < cols ? diff[size_t(r) * (cols + 1) + c] : 0.0;
        for (int off = 1; off < kWave; off <<= 1) {
            const double u = __shfl_up(v, off);
            if (lane >= off) v += u;
        }
        if (c < cols) sol[size_t(c) * rows + r] = direct[size_t(r) * cols + c] + (carry + v);
        carry += __shfl(v, kWave - 1);
    }
}

constexpr int kMaxStaticDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch[kMaxStaticDevices] = {};

// The static scratch of the current device (`dev`), its address taken once
// per device.
int static_scratch(int dev, unsigned char** out) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (dev < 0 || dev >= kMaxStaticDevices) return set_error(MCG_EDEVICE, "device index beyond the scratch table");
    if (!g_scratch[dev]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return set_error(MCG_EDEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch[dev] = static_cast<unsigned char*>(ptr);
    }
    *out = g_scratch[dev];
    return MCG_OK;
}

// mcg_solve with every device buffer in the static scratch:
// [direct rows x cols | diff rows x (cols + 1) | counters | output rows x cols
// | emitter cumulative counts], 8-byte words.
template <typename F>
int run_solve_static(mcg_scene* sc, const HostProblem& p, double* out, mcg_stats* stats) {
    const HostScene& h = sc->host;
    auto t0 = std::chrono::steady_clock::now();
    HIP_TRY(hipSetDevice(sc->device));
    if (p.maxloop <= 0) return set_error(MCG_EINVAL, "maxloop must be positive");
    const int rows = int(p.rows), cols = int(h.cols);
    const size_t nd = size_t(p.rows) * size_t(h.cols), nf = size_t(p.rows) * size_t(h.cols + 1);
    const size_t ncdf = p.emit_cdf.size();
    const size_t words = nd + nf + C_COUNT + nd + ncdf;
    if (p.rows > (1 << 20) || words > MCG_STATIC_SCRATCH_BYTES / 8)
        return set_error(MCG_ERANGE, "the solve needs " + std::to_string(words * 8) +
                                         " bytes of device scratch; mcg_solve_static has " +
                                         std::to_string(size_t(MCG_STATIC_SCRATCH_BYTES)));
    unsigned char* base = nullptr;
    int rc = static_scratch(sc->device, &base);
    if (rc) return rc;
    double* const direct = reinterpret_cast<double*>(base);
    double* const diff = direct + nd;
    unsigned long long* const ctr = reinterpret_cast<unsigned long long*>(diff + nf);
    double* const sol = reinterpret_cast<double*>(ctr + C_COUNT);
    int64_t* const cdf = reinterpret_cast<int64_t*>(sol + nd);
    HIP_TRY(hipMemcpy(cdf, p.emit_cdf.data(), ncdf * sizeof(int64_t), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(zero_kernel, dim3(int(std::min<size_t>(1024, (nd + nf + 255) / 256 + 1))), dim3(256), 0, 0,
                       direct, nd + nf, static_cast<double*>(nullptr), size_t(0), ctr, size_t(C_COUNT));
    HIP_TRY(hipGetLastError());
    KParams<F> K{};
    fill_params<F>(sc, p, K);
    K.emit_cdf = cdf;
    K.direct = direct;
    K.diff = diff;
    K.counters = ctr;
    K.next = ctr + C_NEXT;
    const auto tk = std::chrono::steady_clock::now();
    if ((rc = launch_solve<F>(sc, p, K, 0))) return rc;
    hipLaunchKernelGGL(fold_kernel, dim3(rows), dim3(kWave), 0, 0, direct, diff, rows, cols, sol);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(0));
    const auto tk1 = std::chrono::steady_clock::now();
    unsigned long long hc[C_COUNT];
    HIP_TRY(hipMemcpy(hc, ctr, sizeof(hc), hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(out, sol, nd * sizeof(double), hipMemcpyDeviceToHost));
    normalise_field(h, p, out);
    if (stats) {
        stats->bundles = int64_t(hc[C_BUNDLES]);
        stats->steps = int64_t(hc[C_STEPS]);
        stats->escapes = int64_t(hc[C_ESC]);
        stats->absorbed = int64_t(hc[C_ABS]);
        stats->truncated = int64_t(hc[C_TRUNC]);
        stats->kernel_ms = std::chrono::duration<double, std::milli>(tk1 - tk).count();
        stats->wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    if (int64_t(hc[C_BUNDLES]) != p.count)
        return set_error(MCG_EDEVICE, "kernel finished with bundles unaccounted for");
    return MCG_OK;
}

// ------------------------------------------------------------ event engine host
// The bundle lists of the event engine: kEvGroups regions of `cap` slots,
// one array per field, carved out of one allocation.
template <typename F>
EvState<F> ev_carve(unsigned char* base, int64_t slots) {
    EvState<F> S;
    size_t off = 0;
    auto take = [&](size_t bytes) {
        unsigned char* p = base + off;
        off += (bytes * size_t(slots) + 255) & ~size_t(255);
        return p;
    };
    S.px = reinterpret_cast<F*>(take(sizeof(F)));
    S.py = reinterpret_cast<F*>(take(sizeof(F)));
    S.pz = reinterpret_cast<F*>(take(sizeof(F)));
    S.ux = reinterpret_cast<F*>(take(sizeof(F)));
    S.uy = reinterpret_cast<F*>(take(sizeof(F)));
    S.uz = reinterpret_cast<F*>(take(sizeof(F)));
    S.sn = reinterpret_cast<F*>(take(sizeof(F)));
    S.mode = reinterpret_cast<uint32_t*>(take(4));
    S.sds = reinterpret_cast<int32_t*>(take(4));
    S.nscat = reinterpret_cast<uint32_t*>(take(4));
    S.nbdry = reinterpret_cast<uint32_t*>(take(4));
    S.nstep = reinterpret_cast<uint32_t*>(take(4));
    S.id = reinterpret_cast<int64_t*>(take(8));
    S.hit = reinterpret_cast<int32_t*>(take(4));
    return S;
}

template <typename F>
size_t ev_list_bytes(int64_t slots) {
    size_t total = 0;
    const size_t sizes[] = {sizeof(F), sizeof(F), sizeof(F), sizeof(F), sizeof(F), sizeof(F), sizeof(F),
                            4, 4, 4, 4, 4, 8, 4};
    for (size_t b : sizes) total += (b * size_t(slots) + 255) & ~size_t(255);
    return total;
}

template <typename F>
const void* ev_kernel(int which, bool tally) {
    switch (which) {
        case 0: return tally ? reinterpret_cast<const void*>(ev_round_kernel<F, true>)
                             : reinterpret_cast<const void*>(ev_round_kernel<F, false>);
        case 1: return tally ? reinterpret_cast<const void*>(ev_emit_kernel<F, true>)
                             : reinterpret_cast<const void*>(ev_emit_kernel<F, false>);
        default: return tally ? reinterpret_cast<const void*>(ev_drain_kernel<F, true>)
                              : reinterpret_cast<const void*>(ev_drain_kernel<F, false>);
    }
}

// mcg_solve by the event engine (tune.engine = MCG_ENGINE_EVENT).
template <typename F>
int run_solve_event(mcg_scene* sc, const HostProblem& p, double* out, mcg_counts* counts, mcg_stats* stats) {
    const HostScene& h = sc->host;
    auto t0 = std::chrono::steady_clock::now();
    HIP_TRY(hipSetDevice(sc->device));
    const int rows = int(p.rows);
    const int cols = int(h.cols);
    if (p.rows > (1 << 20) || h.cols * p.rows > (1ll << 31)) return set_error(MCG_ERANGE, "tally too large");
    if (p.maxloop <= 0) return set_error(MCG_EINVAL, "maxloop must be positive");
    const size_t nd = size_t(rows) * cols, nf = size_t(rows) * (cols + 1);
    const size_t nall = nd + nf + C_COUNT;
    int rc;
    if ((rc = sc->ws_tally.reserve(nall))) return rc;
    if (sc->ws_cdf.p == nullptr || sc->ws_cdf.n < p.emit_cdf.size() || sc->ws_cdf_host != p.emit_cdf) {
        if ((rc = sc->ws_cdf.reserve(p.emit_cdf.size()))) return rc;
        HIP_TRY(hipMemcpy(sc->ws_cdf.p, p.emit_cdf.data(), p.emit_cdf.size() * sizeof(int64_t),
                          hipMemcpyHostToDevice));
        sc->ws_cdf_host = p.emit_cdf;
    }
    if (sc->pinned_n < nall) {
        if (sc->pinned) HIP_TRY(hipHostFree(sc->pinned));
        sc->pinned = nullptr;
        sc->pinned_n = 0;
        HIP_TRY(hipHostMalloc(reinterpret_cast<void**>(&sc->pinned), nall * sizeof(double), hipHostMallocDefault));
        sc->pinned_n = nall;
    }
    // bundles in flight and the lists holding them
    const int64_t live = std::max<int64_t>(1, std::min<int64_t>(p.count, sc->tune.live > 0 ? sc->tune.live
                                                                                          : (int64_t(1) << 21)));
    const int64_t cap = live;
    const int64_t slots = int64_t(kEvGroups) * cap;
    const size_t list_bytes = ev_list_bytes<F>(slots);
    if ((rc = sc->ev_state.reserve(2 * list_bytes))) return rc;
    // counters, each on its own cache line: list A (kEvGroups), list B
    // (kEvGroups), then emit (2) and the next id (1) together
    const size_t nctl = size_t(2 * kEvGroups + 1) * kEvPad;
    if ((rc = sc->ev_ctl.reserve(nctl))) return rc;
    unsigned long long* const cA = sc->ev_ctl.p;
    unsigned long long* const cB = cA + kEvGroups * kEvPad;
    unsigned long long* const emitw = cB + kEvGroups * kEvPad;
    unsigned long long* const next_id = emitw + 2;
    DevBuf<uint32_t> cnt;
    const bool want_counts = counts && counts->nscat && counts->nbdry && counts->nstep && counts->fate;
    if (want_counts && (rc = cnt.alloc(size_t(4) * p.count))) return rc;

    double* const direct = sc->ws_tally.p;
    double* const diff = direct + nd;
    unsigned long long* const ctr = reinterpret_cast<unsigned long long*>(direct + nd + nf);
    hipLaunchKernelGGL(zero_kernel, dim3(int(std::min<size_t>(1024, (nd + nf + 255) / 256 + 1))), dim3(256), 0, 0,
                       direct, nd + nf, static_cast<double*>(nullptr), size_t(0), ctr, size_t(C_COUNT));
    hipLaunchKernelGGL(zero_kernel, dim3(1), dim3(256), 0, 0, static_cast<double*>(nullptr), size_t(0),
                       static_cast<double*>(nullptr), size_t(0), sc->ev_ctl.p, nctl);
    HIP_TRY(hipGetLastError());

    EvParams<F> E{};
    fill_params<F>(sc, p, E.K);
    E.K.emit_cdf = sc->ws_cdf.p;
    E.K.direct = direct;
    E.K.diff = diff;
    E.K.counters = ctr;
    E.K.next = ctr + C_NEXT;
    if (want_counts) {
        E.K.c_nscat = cnt.p;
        E.K.c_nbdry = cnt.p + p.count;
        E.K.c_nstep = cnt.p + 2 * p.count;
        E.K.c_fate = cnt.p + 3 * p.count;
    }
    E.cap = cap;
    E.live_target = live;
    EvState<F> SA = ev_carve<F>(sc->ev_state.p, slots), SB = ev_carve<F>(sc->ev_state.p + list_bytes, slots);
    // LDS: geometry, and the tally when it fits (the tables stay in global
    // memory: every launch stages what it uses)
    const size_t geo = lds_bytes<F>(h, rows, false, false);
    // (256 B of the LDS are the appends' static workspace)
    const LdsPlan lp = plan_lds(geo, 0, lds_bytes<F>(h, rows, false, true) - geo, h.has_grid, false,
                                sc->lds_limit - 256);
    if (lp.bytes > sc->lds_limit) return set_error(MCG_ERANGE, "geometry exceeds the LDS budget");
    E.K.lds_tally = lp.tally ? 1 : 0;
    int per_cu = std::max<int>(1, int(sc->lds_limit / std::max<size_t>(lp.bytes, 1)));
    // one workgroup per CU (a large LDS tally): give it the CU's 16 waves
    const int block = sc->tune.block > 0 ? sc->tune.block : (per_cu == 1 ? 1024 : 256);
    per_cu = std::min(per_cu, std::max(1, 2048 / block));
    if (sc->tune.blocks_per_cu > 0) per_cu = std::min(per_cu, int(sc->tune.blocks_per_cu));
    const int64_t max_grid = int64_t(sc->num_cu) * per_cu;
    const int per_thread = sc->tune.per_thread > 0 ? sc->tune.per_thread : 1;
    const int compact_every = sc->tune.compact_every > 0 ? sc->tune.compact_every : 1;
    for (int w = 0; w < 3; ++w)
        HIP_TRY(hipFuncSetAttribute(ev_kernel<F>(w, lp.tally), hipFuncAttributeMaxDynamicSharedMemorySize,
                                    int(lp.bytes)));
    auto grid_for = [&](int64_t n) {
        const int64_t want = (n + int64_t(block) * per_thread - 1) / (int64_t(block) * per_thread);
        return int(std::max<int64_t>(1, std::min<int64_t>(want, max_grid)));
    };
    auto launch = [&](int which, const EvParams<F>& P, int grid) -> int {
        EvParams<F> P2 = P;
        void* args[] = {&P2};
        HIP_TRY(hipLaunchKernel(ev_kernel<F>(which, lp.tally), dim3(grid), dim3(block), args, lp.bytes, 0));
        HIP_TRY(hipGetLastError());
        return MCG_OK;
    };
    if (!sc->ev0) HIP_TRY(hipEventCreate(&sc->ev0));
    if (!sc->ev1) HIP_TRY(hipEventCreate(&sc->ev1));
    HIP_TRY(hipEventRecord(sc->ev0, 0));
    // A holds the current list; the first emission fills it
    E.in = SB;
    E.out = SA;
    E.in_count = cB;
    E.out_count = cA;
    E.emit = emitw;
    hipLaunchKernelGGL(ev_prep_kernel, dim3(1), dim3(64), 0, 0, cA, cB, next_id, emitw, p.count, live);
    if ((rc = launch(1, E, grid_for(live)))) return rc;
    bool a_current = true;
    int64_t rounds = 0;
    const int64_t drain_at = std::max<int64_t>(int64_t(sc->num_cu) * 256, live / 16);
    for (;;) {
        // a batch of rounds, then one look at the lists
        for (int r = 0; r < 8; ++r, ++rounds) {
            EvParams<F> R = E;
            R.in = a_current ? SA : SB;
            R.out = a_current ? SB : SA;
            R.in_count = a_current ? cA : cB;
            R.out_count = a_current ? cB : cA;
            if ((rounds + 1) % compact_every != 0) {
                // in place: holes stay until the next regrouping
                R.regroup = 0;
                if ((rc = launch(0, R, grid_for(live)))) return rc;
                continue;
            }
            R.regroup = 1;
            if ((rc = launch(0, R, grid_for(live)))) return rc;
            // emission up to the live target into the new list; the old
            // list's counters are cleared for the round after
            hipLaunchKernelGGL(ev_prep_kernel, dim3(1), dim3(64), 0, 0, R.out_count, R.in_count, next_id, emitw,
                               p.count, live);
            if ((rc = launch(1, R, grid_for(live)))) return rc;
            a_current = !a_current;
        }
        std::vector<unsigned long long> hc(nctl);
        HIP_TRY(hipMemcpy(hc.data(), sc->ev_ctl.p, nctl * sizeof(unsigned long long), hipMemcpyDeviceToHost));
        const unsigned long long* cur = hc.data() + (a_current ? 0 : kEvGroups * kEvPad);
        int64_t in_list = 0;
        for (int g = 0; g < kEvGroups; ++g) in_list += int64_t(cur[g * kEvPad]);
        const bool emitted_all = int64_t(hc[size_t(2 * kEvGroups) * kEvPad + 2]) >= p.count;
        if (emitted_all && in_list <= drain_at) {
            if (in_list > 0) {
                EvParams<F> R = E;
                R.in = a_current ? SA : SB;
                R.in_count = a_current ? cA : cB;
                if ((rc = launch(2, R, grid_for(in_list)))) return rc;
            }
            break;
        }
        // every round takes each live bundle one step, so a solve ends within
        // maxloop rounds of its last emission; more means a fault
        if (rounds > std::min<int64_t>(p.maxloop, int64_t(1) << 20) + (p.count / live + 1) * 8 + 64)
            return set_error(MCG_EDEVICE, "event engine did not finish within maxloop rounds");
    }
    HIP_TRY(hipEventRecord(sc->ev1, 0));
    HIP_TRY(hipMemcpyAsync(sc->pinned, direct, nall * sizeof(double), hipMemcpyDeviceToHost, 0));
    HIP_TRY(hipStreamSynchronize(0));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, sc->ev0, sc->ev1));
    unsigned long long hc[C_COUNT];
    std::memcpy(hc, sc->pinned + nd + nf, sizeof(hc));
    if (want_counts) {
        HIP_TRY(hipMemcpy(counts->nscat, cnt.p, p.count * 4, hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(counts->nbdry, cnt.p + p.count, p.count * 4, hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(counts->nstep, cnt.p + 2 * p.count, p.count * 4, hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(counts->fate, cnt.p + 3 * p.count, p.count * 4, hipMemcpyDeviceToHost));
    }
    finalise_field(h, p, sc->pinned, sc->pinned + nd, out);
    if (stats) {
        stats->bundles = int64_t(hc[C_BUNDLES]);
        stats->steps = int64_t(hc[C_STEPS]);
        stats->escapes = int64_t(hc[C_ESC]);
        stats->absorbed = int64_t(hc[C_ABS]);
        stats->truncated = int64_t(hc[C_TRUNC]);
        stats->kernel_ms = ms;
        stats->wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    if (int64_t(hc[C_BUNDLES]) != p.count)
        return set_error(MCG_EDEVICE, "event engine finished with bundles unaccounted for");
    return MCG_OK;
}

// Host side of mcg_trace_from: buffers for the one lane's tallies (kept on
// the GPU and dropped), its counts and its records.
template <typename F>
int run_trace_from(mcg_scene* sc, const HostProblem& p, const TraceStart<F>& start, double* trace,
                   int64_t cap, int64_t* nrec) {
    const HostScene& h = sc->host;
    HIP_TRY(hipSetDevice(sc->device));
    const int rows = int(p.rows), cols = int(h.cols);
    if (p.rows > (1 << 20) || h.cols * p.rows > (1ll << 31)) return set_error(MCG_ERANGE, "tally too large");
    if (p.maxloop <= 0) return set_error(MCG_EINVAL, "maxloop must be positive");
    const size_t nd = size_t(rows) * cols, nf = size_t(rows) * (cols + 1);
    DevBuf<double> tally, tbuf;
    DevBuf<uint32_t> cnt;
    int rc;
    if ((rc = tally.alloc(nd + nf + 1))) return rc;
    if ((rc = cnt.alloc(4))) return rc;
    if ((rc = tbuf.alloc(size_t(cap) * MCG_TRACE_WIDTH))) return rc;
    HIP_TRY(hipMemset(tally.p, 0, tally.n * sizeof(double)));
    HIP_TRY(hipMemset(cnt.p, 0, cnt.n * sizeof(uint32_t)));
    HIP_TRY(hipMemset(tbuf.p, 0, tbuf.n * sizeof(double)));
    KParams<F> K{};
    fill_params<F>(sc, p, K);
    K.direct = tally.p;
    K.diff = tally.p + nd;
    K.c_nscat = cnt.p;
    K.c_nbdry = cnt.p + 1;
    K.c_nstep = cnt.p + 2;
    K.c_fate = cnt.p + 3;
    K.trace_id = p.first;
    K.trace = tbuf.p;
    K.trace_cap = cap;
    const size_t lds = lds_bytes<F>(h, rows, false, false);
    if (lds > sc->lds_limit) return set_error(MCG_ERANGE, "geometry exceeds the LDS budget");
    const void* fn = reinterpret_cast<const void*>(trace_from_kernel<F>);
    HIP_TRY(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, int(lds)));
    TraceStart<F> s = start;
    void* args[] = {&K, &s};
    HIP_TRY(hipLaunchKernel(fn, dim3(1), dim3(kWave), args, lds, 0));
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipDeviceSynchronize());
    uint32_t hc[4];
    HIP_TRY(hipMemcpy(hc, cnt.p, sizeof(hc), hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(trace, tbuf.p, tbuf.n * sizeof(double), hipMemcpyDeviceToHost));
    *nrec = std::min<int64_t>(int64_t(hc[2]), cap);
    return MCG_OK;
}

// Launch plan of a batch: per-solve tally offsets, LDS layout, launch shape.
struct BatchPlan {
    std::vector<size_t> off_direct, off_diff;
    size_t n_direct = 0, n_diff = 0;
    size_t lds = 0;
    bool tables = true, tally = true;
    int walk = 0;              // axes of deferred tally walks (walk_axes)
    int block = 256, grid = 0, chunk = kChunk;
};

// LDS layout, block size and grid of a batch launch over these solves.
template <typename F>
int plan_shape(int n, mcg_scene* const* scenes, const HostProblem* P, const mcg_tune* tune,
               BatchPlan& B) {
    const mcg_scene* s0 = scenes[0];
    size_t geo = 0, tab = 0, tally_bytes = 0;
    bool grid = false, grid_2d = true;
    int64_t cells = 0;
    for (int i = 0; i < n; ++i) {
        const HostScene& h = scenes[i]->host;
        if (scenes[i]->device != s0->device) return set_error(MCG_EINVAL, "batch scenes on different devices");
        const size_t g = lds_bytes<F>(h, int(P[i].rows), false, false);
        geo = std::max(geo, g);
        tab = std::max(tab, lds_bytes<F>(h, int(P[i].rows), true, false) - g);
        tally_bytes = std::max(tally_bytes, lds_bytes<F>(h, int(P[i].rows), false, true) - g);
        grid = grid || h.has_grid;
        grid_2d = grid_2d && h.grid_2d;
        cells = std::max(cells, h.grid_cells);
    }
    if (tune && (tune->defer_walks < -1 || tune->defer_walks > 1))
        return set_error(MCG_EINVAL, "defer_walks must be -1, 0 or 1");
    const LdsPlan lp = plan_lds(geo, tab, tally_bytes, grid, !(tune && tune->lds_tables < 0), s0->lds_limit);
    B.walk = walk_axes<F>(grid, grid_2d, cells, KernelKind::Batch, tune ? tune->defer_walks : 0);
    B.tables = lp.tables;
    B.tally = lp.tally;
    B.lds = lp.bytes;
    if (B.lds > s0->lds_limit) return set_error(MCG_ERANGE, "geometry exceeds the LDS budget");
    // occupancy: LDS per CU (the per-block limit is the CU's LDS on gfx950)
    // and at most 2048 threads per CU
    int per_cu = std::max<int>(1, int(s0->lds_limit / std::max<size_t>(B.lds, 1)));
    B.block = (tune && tune->block) ? tune->block : (per_cu == 1 ? 1024 : 512);
    if (B.block < 64 || B.block > 1024 || B.block % 64 != 0)
        return set_error(MCG_EINVAL, "block must be a multiple of 64 from 64 to 1024");
    per_cu = std::min(per_cu, std::max(1, 2048 / B.block));
    if (tune && tune->blocks_per_cu > 0) per_cu = std::min(per_cu, int(tune->blocks_per_cu));
    B.grid = s0->num_cu * per_cu;
    B.chunk = (tune && tune->chunk > 0) ? tune->chunk : kChunk;
    return MCG_OK;
}

template <typename F>
int plan_batch(int n, mcg_scene* const* scenes, const HostProblem* P, const mcg_tune* tune,
               BatchPlan& B) {
    int rc = plan_shape<F>(n, scenes, P, tune, B);
    if (rc) return rc;
    B.off_direct.resize(n);
    B.off_diff.resize(n);
    B.n_direct = B.n_diff = 0;
    int64_t total = 0;
    for (int i = 0; i < n; ++i) {
        const HostScene& h = scenes[i]->host;
        B.off_direct[i] = B.n_direct;
        B.off_diff[i] = B.n_diff;
        B.n_direct += size_t(P[i].rows) * h.cols;
        B.n_diff += size_t(P[i].rows) * (h.cols + 1);
        total += P[i].count;
    }
    // no more workgroups than there are bundles to give each lane one
    B.grid = int(std::max<int64_t>(1, std::min<int64_t>(B.grid, (total + B.block - 1) / B.block)));
    return MCG_OK;
}

// Enqueue one batch launch on `stream`.
template <typename F>
int launch_batch(const BatchPlan& B, const KParams<F>* kp, int nsolves, hipStream_t stream) {
    const void* fn = pick_kernel<F>(KernelKind::Batch, B.tables, B.tally, B.walk);
    HIP_TRY(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, int(B.lds)));
    void* args[] = {&kp, &nsolves};
    HIP_TRY(hipLaunchKernel(fn, dim3(std::max(1, B.grid)), dim3(B.block), args, B.lds, stream));
    HIP_TRY(hipGetLastError());
    return MCG_OK;
}

template <typename F>
int run_batch(int n, mcg_scene* const* scenes, const HostProblem* P, double* const* outs,
              mcg_stats* stats, const mcg_tune* tune) {
    auto t0 = std::chrono::steady_clock::now();
    HIP_TRY(hipSetDevice(scenes[0]->device));
    BatchPlan B;
    int rc = plan_batch<F>(n, scenes, P, tune, B);
    if (rc) return rc;
    DevBuf<double> direct, diff;
    DevBuf<unsigned long long> ctr;
    DevBuf<int64_t> cdf;
    DevBuf<KParams<F>> kp;
    std::vector<int64_t> cdf_all;
    std::vector<size_t> cdf_off(n);
    for (int i = 0; i < n; ++i) {
        cdf_off[i] = cdf_all.size();
        cdf_all.insert(cdf_all.end(), P[i].emit_cdf.begin(), P[i].emit_cdf.end());
    }
    if ((rc = direct.alloc(B.n_direct))) return rc;
    if ((rc = diff.alloc(B.n_diff))) return rc;
    if ((rc = ctr.alloc(size_t(C_COUNT) * n))) return rc;
    if ((rc = cdf.upload(cdf_all.data(), cdf_all.size()))) return rc;
    std::vector<KParams<F>> K(n);
    for (int i = 0; i < n; ++i) {
        K[i] = KParams<F>{};
        fill_params<F>(scenes[i], P[i], K[i]);
        K[i].emit_cdf = cdf.p + cdf_off[i];
        K[i].direct = direct.p + B.off_direct[i];
        K[i].diff = diff.p + B.off_diff[i];
        K[i].counters = ctr.p + size_t(C_COUNT) * i;
        K[i].next = K[i].counters + C_NEXT;
        K[i].chunk = B.chunk;
        K[i].lds_tally = B.tally ? 1 : 0;
    }
    if ((rc = kp.upload(K.data(), K.size()))) return rc;
    HIP_TRY(hipMemset(direct.p, 0, direct.n * sizeof(double)));
    HIP_TRY(hipMemset(diff.p, 0, diff.n * sizeof(double)));
    HIP_TRY(hipMemset(ctr.p, 0, ctr.n * sizeof(unsigned long long)));
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    HIP_TRY(hipEventRecord(e0, 0));
    if ((rc = launch_batch<F>(B, kp.p, n, 0))) return rc;
    HIP_TRY(hipEventRecord(e1, 0));
    HIP_TRY(hipEventSynchronize(e1));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    std::vector<double> hd(direct.n), hf(diff.n);
    std::vector<unsigned long long> hc(ctr.n);
    HIP_TRY(hipMemcpy(hd.data(), direct.p, hd.size() * sizeof(double), hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(hf.data(), diff.p, hf.size() * sizeof(double), hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(hc.data(), ctr.p, hc.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost));
    const double wall = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    for (int i = 0; i < n; ++i) {
        finalise_field(scenes[i]->host, P[i], hd.data() + B.off_direct[i], hf.data() + B.off_diff[i], outs[i]);
        const unsigned long long* c = hc.data() + size_t(C_COUNT) * i;
        if (stats) {
            stats[i].bundles = int64_t(c[C_BUNDLES]);
            stats[i].steps = int64_t(c[C_STEPS]);
            stats[i].escapes = int64_t(c[C_ESC]);
            stats[i].absorbed = int64_t(c[C_ABS]);
            stats[i].truncated = int64_t(c[C_TRUNC]);
            stats[i].kernel_ms = ms;
            stats[i].wall_ms = wall;
        }
        if (int64_t(c[C_BUNDLES]) != P[i].count)
            return set_error(MCG_EDEVICE, "batch finished with bundles unaccounted for");
    }
    return MCG_OK;
}

}  // namespace
}  // namespace mcg


namespace mcg {
// Device side of a scene whose host part is built: query the current GPU and
// upload geometry (both precisions), tables and Inter candidates.
int scene_upload(mcg_scene* sc) {
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev == 0) return set_error(MCG_EDEVICE, "no HIP device available");
    HIP_TRY(hipGetDevice(&sc->device));
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, sc->device));
    sc->num_cu = prop.multiProcessorCount;
    sc->lds_limit = prop.sharedMemPerBlock;
    int rc;
    if ((rc = upload_scene(sc->host, sc->d64))) return rc;
    if ((rc = upload_scene(sc->host, sc->d32))) return rc;
    if ((rc = sc->scat_alias.upload(sc->host.scat_alias.data(), sc->host.scat_alias.size()))) return rc;
    if ((rc = sc->flux_alias.upload(sc->host.flux_alias.data(), sc->host.flux_alias.size()))) return rc;
    if ((rc = sc->inter_list.upload(sc->host.inter_list.data(), sc->host.inter_list.size()))) return rc;
    return MCG_OK;
}

// Device bytes a scene holds once uploaded.
size_t scene_bytes(const HostScene& h) {
    const size_t modes = h.mode_cell.size();
    return h.planes.size() * (sizeof(DevPlane<double>) + sizeof(DevPlane<float>)) +
           h.sdoms.size() * (sizeof(DevSdom<double>) + sizeof(DevSdom<float>)) +
           h.emitters.size() * (sizeof(DevEmitter<double>) + sizeof(DevEmitter<float>)) +
           h.pores.size() * (sizeof(DevPore<double>) + sizeof(DevPore<float>)) +
           modes * (2 * 8 + 2 * 4 + 2 * 8) + h.inter_list.size() * 4;
}

// One device block holding everything a served request needs: the scene in
// both precisions, the alias tables, the Inter list and the emitter counts
// (uploaded from a packed host image in one copy), then the tallies and
// counters (cleared on the device), contiguous so that one copy reads them
// back.  Offsets in bytes, each buffer 256-byte aligned.
struct ImageLayout {
    size_t planes64, sdoms64, emit64, pores64, vel64, mfp64;
    size_t planes32, sdoms32, emit32, pores32, vel32, mfp32;
    size_t scat, flux, inter, cdf;
    size_t upload;   // bytes of the image copied to the device
    size_t direct, nd, nf;   // direct (nd doubles), then diff (nf doubles), then the counters
    size_t total;
};

namespace {
template <typename T>
void put(std::vector<unsigned char>& img, size_t off, const T* src, size_t count) {
    if (count) std::memcpy(img.data() + off, src, count * sizeof(T));
}

template <typename F>
void pack_precision(const HostScene& h, std::vector<unsigned char>& img, size_t planes, size_t sdoms,
                    size_t emit, size_t pores, size_t vel, size_t mfp) {
    const double eps_scale = sizeof(F) == 4 ? double(FLT_EPSILON) / DBL_EPSILON : 1.0;
    std::vector<DevPlane<F>> pl(h.planes.size());
    for (size_t i = 0; i < pl.size(); ++i) convert_plane(h.planes[i], pl[i]);
    std::vector<DevSdom<F>> sd(h.sdoms.size());
    for (size_t i = 0; i < sd.size(); ++i) convert_sdom(h.sdoms[i], sd[i], eps_scale);
    std::vector<DevEmitter<F>> em(h.emitters.size());
    for (size_t i = 0; i < em.size(); ++i) convert_emitter(h.emitters[i], em[i]);
    std::vector<DevPore<F>> po(h.pores.size());
    for (size_t i = 0; i < po.size(); ++i) convert_pore(h.pores[i], po[i], eps_scale);
    std::vector<F> mv(h.mode_vel.size()), mm(h.mode_mfp.size());
    for (size_t i = 0; i < mv.size(); ++i) { mv[i] = F(h.mode_vel[i]); mm[i] = F(h.mode_mfp[i]); }
    put(img, planes, pl.data(), pl.size());
    put(img, sdoms, sd.data(), sd.size());
    put(img, emit, em.data(), em.size());
    put(img, pores, po.data(), po.size());
    put(img, vel, mv.data(), mv.size());
    put(img, mfp, mm.data(), mm.size());
}
}  // namespace

// Lays out and packs a request's device image (host work only).
void pack_request(const HostScene& h, const HostProblem& P, ImageLayout& L, std::vector<unsigned char>& img) {
    size_t off = 0;
    auto take = [&](size_t bytes) {
        const size_t at = off;
        off += (bytes + 255) & ~size_t(255);
        return at;
    };
    const size_t np = h.planes.size(), ns = h.sdoms.size(), ne = h.emitters.size(), npo = h.pores.size();
    const size_t nm = h.mode_vel.size();
    L.planes64 = take(np * sizeof(DevPlane<double>));
    L.sdoms64 = take(ns * sizeof(DevSdom<double>));
    L.emit64 = take(ne * sizeof(DevEmitter<double>));
    L.pores64 = take(npo * sizeof(DevPore<double>));
    L.vel64 = take(nm * sizeof(double));
    L.mfp64 = take(nm * sizeof(double));
    L.planes32 = take(np * sizeof(DevPlane<float>));
    L.sdoms32 = take(ns * sizeof(DevSdom<float>));
    L.emit32 = take(ne * sizeof(DevEmitter<float>));
    L.pores32 = take(npo * sizeof(DevPore<float>));
    L.vel32 = take(nm * sizeof(float));
    L.mfp32 = take(nm * sizeof(float));
    L.scat = take(h.scat_alias.size() * sizeof(uint64_t));
    L.flux = take(h.flux_alias.size() * sizeof(uint64_t));
    L.inter = take(h.inter_list.size() * sizeof(int32_t));
    L.cdf = take(P.emit_cdf.size() * sizeof(int64_t));
    L.upload = off;
    L.nd = size_t(P.rows) * size_t(h.cols);
    L.nf = size_t(P.rows) * size_t(h.cols + 1);
    L.direct = take((L.nd + L.nf + C_COUNT) * sizeof(double));
    L.total = off;
    img.assign(L.upload, 0);
    pack_precision<double>(h, img, L.planes64, L.sdoms64, L.emit64, L.pores64, L.vel64, L.mfp64);
    pack_precision<float>(h, img, L.planes32, L.sdoms32, L.emit32, L.pores32, L.vel32, L.mfp32);
    put(img, L.scat, h.scat_alias.data(), h.scat_alias.size());
    put(img, L.flux, h.flux_alias.data(), h.flux_alias.size());
    put(img, L.inter, h.inter_list.data(), h.inter_list.size());
    put(img, L.cdf, P.emit_cdf.data(), P.emit_cdf.size());
}

// Points the scene's device buffers into a request's block.
void view_scene(mcg_scene* sc, unsigned char* b, const ImageLayout& L) {
    const HostScene& h = sc->host;
    const size_t np = h.planes.size(), ns = h.sdoms.size(), ne = h.emitters.size(), npo = h.pores.size();
    const size_t nm = h.mode_vel.size();
    sc->d64.planes.view(reinterpret_cast<DevPlane<double>*>(b + L.planes64), np);
    sc->d64.sdoms.view(reinterpret_cast<DevSdom<double>*>(b + L.sdoms64), ns);
    sc->d64.emitters.view(reinterpret_cast<DevEmitter<double>*>(b + L.emit64), ne);
    sc->d64.pores.view(reinterpret_cast<DevPore<double>*>(b + L.pores64), npo);
    sc->d64.mode_vel.view(reinterpret_cast<double*>(b + L.vel64), nm);
    sc->d64.mode_mfp.view(reinterpret_cast<double*>(b + L.mfp64), nm);
    sc->d32.planes.view(reinterpret_cast<DevPlane<float>*>(b + L.planes32), np);
    sc->d32.sdoms.view(reinterpret_cast<DevSdom<float>*>(b + L.sdoms32), ns);
    sc->d32.emitters.view(reinterpret_cast<DevEmitter<float>*>(b + L.emit32), ne);
    sc->d32.pores.view(reinterpret_cast<DevPore<float>*>(b + L.pores32), npo);
    sc->d32.mode_vel.view(reinterpret_cast<float*>(b + L.vel32), nm);
    sc->d32.mode_mfp.view(reinterpret_cast<float*>(b + L.mfp32), nm);
    sc->scat_alias.view(reinterpret_cast<uint64_t*>(b + L.scat), h.scat_alias.size());
    sc->flux_alias.view(reinterpret_cast<uint64_t*>(b + L.flux), h.flux_alias.size());
    sc->inter_list.view(reinterpret_cast<int32_t*>(b + L.inter), h.inter_list.size());
}
}  // namespace mcg


// ================================================================ graph batch
namespace mcg {

struct GraphBatchBase {
    virtual ~GraphBatchBase() = default;
    virtual int run(const uint64_t* seeds, double* const* outs, mcg_stats* stats) = 0;
    virtual int64_t need(int i) const = 0;
    int n = 0;
};

template <typename F>
struct GraphBatch : GraphBatchBase {
    std::vector<mcg_scene*> sc;
    std::vector<HostProblem> P;
    BatchPlan B;
    DevBuf<double> direct, diff;
    DevBuf<unsigned long long> ctr;
    DevBuf<int64_t> cdf;
    DevBuf<KParams<F>> kp;
    std::vector<KParams<F>> K;
    hipStream_t stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;

    ~GraphBatch() override {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        if (stream) (void)hipStreamDestroy(stream);
    }
    int64_t need(int i) const override { return P[i].rows * sc[i]->host.cols; }

    int create(const mcg_tune* tune) {
        HIP_TRY(hipSetDevice(sc[0]->device));
        int rc = plan_batch<F>(n, sc.data(), P.data(), tune, B);
        if (rc) return rc;
        std::vector<int64_t> cdf_all;
        std::vector<size_t> cdf_off(n);
        for (int i = 0; i < n; ++i) {
            cdf_off[i] = cdf_all.size();
            cdf_all.insert(cdf_all.end(), P[i].emit_cdf.begin(), P[i].emit_cdf.end());
        }
        if ((rc = direct.alloc(B.n_direct))) return rc;
        if ((rc = diff.alloc(B.n_diff))) return rc;
        if ((rc = ctr.alloc(size_t(C_COUNT) * n))) return rc;
        if ((rc = cdf.upload(cdf_all.data(), cdf_all.size()))) return rc;
        K.resize(n);
        for (int i = 0; i < n; ++i) {
            K[i] = KParams<F>{};
            fill_params<F>(sc[i], P[i], K[i]);
            K[i].emit_cdf = cdf.p + cdf_off[i];
            K[i].direct = direct.p + B.off_direct[i];
            K[i].diff = diff.p + B.off_diff[i];
            K[i].counters = ctr.p + size_t(C_COUNT) * i;
            K[i].next = K[i].counters + C_NEXT;
            K[i].chunk = B.chunk;
            K[i].lds_tally = B.tally ? 1 : 0;
        }
        if ((rc = kp.upload(K.data(), K.size()))) return rc;
        HIP_TRY(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        HIP_TRY(hipEventCreate(&e0));
        HIP_TRY(hipEventCreate(&e1));
        // the kernel's LDS limit is an attribute of the function, set before capture
        HIP_TRY(hipFuncSetAttribute(pick_kernel<F>(KernelKind::Batch, B.tables, B.tally, B.walk),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, int(B.lds)));
        HIP_TRY(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        const size_t total = direct.n + diff.n + ctr.n;
        const int zblocks = int(std::min<size_t>(1024, (total + 255) / 256));
        hipLaunchKernelGGL(zero_kernel, dim3(std::max(1, zblocks)), dim3(256), 0, stream, direct.p, direct.n,
                           diff.p, diff.n, ctr.p, ctr.n);
        rc = launch_batch<F>(B, kp.p, n, stream);
        hipGraph_t g = nullptr;
        const hipError_t ec = hipStreamEndCapture(stream, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (ec != hipSuccess) return set_error(MCG_EDEVICE, std::string("graph capture: ") + hipGetErrorString(ec));
        graph = g;
        HIP_TRY(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
        return MCG_OK;
    }

    int run(const uint64_t* seeds, double* const* outs, mcg_stats* stats) override {
        auto t0 = std::chrono::steady_clock::now();
        HIP_TRY(hipSetDevice(sc[0]->device));
        if (seeds) {
            for (int i = 0; i < n; ++i) {
                K[i].k0 = uint32_t(seeds[i]);
                K[i].k1 = uint32_t(seeds[i] >> 32);
            }
            HIP_TRY(hipMemcpy(kp.p, K.data(), K.size() * sizeof(KParams<F>), hipMemcpyHostToDevice));
        }
        HIP_TRY(hipEventRecord(e0, stream));
        HIP_TRY(hipGraphLaunch(exec, stream));
        HIP_TRY(hipEventRecord(e1, stream));
        HIP_TRY(hipEventSynchronize(e1));
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
        std::vector<double> hd(direct.n), hf(diff.n);
        std::vector<unsigned long long> hc(ctr.n);
        HIP_TRY(hipMemcpy(hd.data(), direct.p, hd.size() * sizeof(double), hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(hf.data(), diff.p, hf.size() * sizeof(double), hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(hc.data(), ctr.p, hc.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost));
        const double wall = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        for (int i = 0; i < n; ++i) {
            finalise_field(sc[i]->host, P[i], hd.data() + B.off_direct[i], hf.data() + B.off_diff[i], outs[i]);
            const unsigned long long* c = hc.data() + size_t(C_COUNT) * i;
            if (stats) {
                stats[i].bundles = int64_t(c[C_BUNDLES]);
                stats[i].steps = int64_t(c[C_STEPS]);
                stats[i].escapes = int64_t(c[C_ESC]);
                stats[i].absorbed = int64_t(c[C_ABS]);
                stats[i].truncated = int64_t(c[C_TRUNC]);
                stats[i].kernel_ms = ms;
                stats[i].wall_ms = wall;
            }
            if (int64_t(c[C_BUNDLES]) != P[i].count)
                return set_error(MCG_EDEVICE, "batch finished with bundles unaccounted for");
        }
        return MCG_OK;
    }
};

}  // namespace mcg

// ====================================================================== serving
// Continuous batching: requests are admitted as device memory allows, their
// tallies and counters stay on the GPU across steps, and every step hands
// each active request a fair share of a per-step bundle budget and runs all
// the shares in one persistent batch launch per precision.  Bundles never
// straddle steps: a share runs its bundles to termination inside the step.
//
// Steps are pipelined: a call plans and enqueues step k+1 (bundle cursors
// advance when work is handed out), then waits for step k and retires the
// requests it finished, so the GPU runs step k+1 while the host copies
// results, admits arrivals and plans.  Consecutive steps go to two streams
// of the default priority, so step k+1's launch starts on the CUs that step
// k's last, long-lived bundles leave idle instead of waiting for the whole
// of step k.  Both launches add into the tallies of a request they share
// (FP64 atomics); a request retires only once every step holding its
// bundles has finished.  The streams do not synchronise with the null
// stream; device memory of retired requests is freed only when the server
// is idle, because freeing can stall the device.
namespace mcg {

struct Request {
    int64_t id = 0;
    std::unique_ptr<mcg_scene> scene;
    HostProblem P;
    int64_t cursor = 0;     // bundles handed out (offset from P.first)
    int64_t done = 0;       // bundles of completed steps
    size_t bytes = 0;       // device memory held while admitted (the block)
    DevBuf<double> direct, diff;
    DevBuf<unsigned long long> ctr;
    DevBuf<int64_t> cdf;
    int state = 0;          // 0 queued, 1 active, 2 done, 3 failed
    std::vector<double> result;
    mcg_stats stats{};
    std::chrono::steady_clock::time_point t_submit, t_done;
    // built by a worker thread after submission: the material (from copies
    // of the spec's arrays), the problem and the packed device image
    std::vector<double> disp, relax;
    mcg_material_spec mspec{};
    mcg_problem_spec pspec{};
    ImageLayout lay{};
    std::vector<unsigned char> image;
    int built = 0;          // 0 building, 1 built, -1 failed (guarded by the server's mutex)
    int err_code = 0;
    std::string err_msg;
    void* block = nullptr;  // the device block, while admitted and until buried
};

// One enqueued step: its launches' buffers, timing events and hand-outs.
struct StepSlot {
    DevBuf<KParams<float>> kp32;
    DevBuf<KParams<double>> kp64;
    DevBuf<unsigned long long> nx32, nx64;   // per-request bundle claim counters of the step
    hipStream_t stream = nullptr;            // this slot's stream (default priority)
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipEvent_t ez = nullptr;                 // after the step's clearing of newly admitted requests
    bool ez_set = false;
    bool live = false;
    std::vector<std::pair<Request*, int64_t>> handed;
    ~StepSlot() {
        if (stream) {
            (void)hipStreamSynchronize(stream);
            (void)hipStreamDestroy(stream);
        }
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        if (ez) (void)hipEventDestroy(ez);
    }
};

}  // namespace mcg

struct mcg_server {
    size_t budget = 0, used = 0;
    int device = 0, num_cu = 0;
    size_t lds_limit = 0;
    // host threads building submitted requests (material, problem, image)
    std::mutex mu;
    std::condition_variable cv_todo, cv_built;
    std::deque<mcg::Request*> todo;
    std::vector<std::thread> workers;
    bool stopping = false;
    int64_t step_bundles = 0;
    mcg_tune tune{};
    int64_t next_id = 1;
    std::deque<mcg::Request*> queue;
    std::vector<mcg::Request*> active;
    std::map<int64_t, std::unique_ptr<mcg::Request>> all;
    std::vector<std::unique_ptr<mcg::Request>> graveyard;   // retired, device memory not yet freed
    std::vector<int64_t> completed;                          // ids retired by the last step call
    std::vector<mcg::Request*> fresh;                        // admitted, tallies not yet cleared
    mcg::StepSlot slot[2];
    int cur = 0;                                             // slot of the step in flight
    double kernel_ms_total = 0.0;
    int64_t steps_total = 0;
};

namespace mcg {
namespace {

// Host work of a submitted request, on a worker thread: the material tables,
// the problem and the packed device image (no HIP call).
int build_request(Request* r) {
    int rc = build_material(&r->mspec, r->scene->host);
    if (!rc) rc = build_problem(&r->pspec, r->scene->host, r->P);
    if (rc) return rc;
    pack_request(r->scene->host, r->P, r->lay, r->image);
    r->bytes = r->lay.total;
    return MCG_OK;
}

void build_worker(mcg_server* sv) {
    for (;;) {
        Request* r = nullptr;
        {
            std::unique_lock<std::mutex> lk(sv->mu);
            sv->cv_todo.wait(lk, [&] { return sv->stopping || !sv->todo.empty(); });
            if (sv->stopping) return;   // the server is going away: what is left is not built
            r = sv->todo.front();
            sv->todo.pop_front();
        }
        int rc;
        try {
            rc = build_request(r);
        } catch (const std::bad_alloc&) {   // nothing may escape a build thread
            rc = set_error(MCG_ENOMEM, "out of host memory building a request");
        } catch (const std::exception& e) {
            rc = set_error(MCG_EINVAL, std::string("building a request: ") + e.what());
        }
        {
            std::lock_guard<std::mutex> lk(sv->mu);
            r->built = rc ? -1 : 1;
            if (rc) {
                r->err_code = rc;
                r->err_msg = t_err;
            }
        }
        sv->cv_built.notify_all();
    }
}

// Blocks until the request's build has finished (or failed).
void wait_built(mcg_server* sv, Request* r) {
    std::unique_lock<std::mutex> lk(sv->mu);
    sv->cv_built.wait(lk, [&] { return r->built != 0; });
}

void stop_workers(mcg_server* sv) {
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        sv->stopping = true;
    }
    sv->cv_todo.notify_all();
    for (std::thread& t : sv->workers)
        if (t.joinable()) t.join();
    sv->workers.clear();
}

// Free device memory of retired requests (called when nothing is in flight).
void bury(mcg_server* sv) {
    for (auto& r : sv->graveyard) {
        r->scene.reset();
        if (r->block) (void)hipFree(r->block);
        r->block = nullptr;
    }
    sv->graveyard.clear();
}

// A request's device side: one block, its image uploaded in one copy.
int admit(mcg_server* sv, Request* r) {
    const ImageLayout& L = r->lay;
    void* b = nullptr;
    hipError_t e = hipMalloc(&b, L.total);
    if (e != hipSuccess && !sv->graveyard.empty()) {
        // retired requests' blocks are freed lazily: free them now, then retry
        (void)hipGetLastError();
        bury(sv);
        e = hipMalloc(&b, L.total);
    }
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return set_error(MCG_ENOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    }
    r->block = b;
    if (L.upload) HIP_TRY(hipMemcpy(b, r->image.data(), L.upload, hipMemcpyHostToDevice));
    std::vector<unsigned char>().swap(r->image);
    unsigned char* const base = static_cast<unsigned char*>(b);
    mcg_scene* sc = r->scene.get();
    sc->device = sv->device;
    sc->num_cu = sv->num_cu;
    sc->lds_limit = sv->lds_limit;
    view_scene(sc, base, L);
    double* const direct = reinterpret_cast<double*>(base + L.direct);
    r->direct.view(direct, L.nd);
    r->diff.view(direct + L.nd, L.nf);
    r->ctr.view(reinterpret_cast<unsigned long long*>(direct + L.nd + L.nf), C_COUNT);
    r->cdf.view(reinterpret_cast<int64_t*>(base + L.cdf), r->P.emit_cdf.size());
    // cleared on the stream of the first step that uses the request
    sv->fresh.push_back(r);
    r->state = 1;
    sv->used += r->bytes;
    sv->active.push_back(r);
    return MCG_OK;
}

int retire(mcg_server* sv, Request* r) {
    // tallies and counters are contiguous in the block: one copy
    const size_t nd = r->direct.n, nf = r->diff.n;
    std::vector<double> h(nd + nf + C_COUNT);
    HIP_TRY(hipMemcpy(h.data(), r->direct.p, h.size() * sizeof(double), hipMemcpyDeviceToHost));
    unsigned long long hc[C_COUNT];
    std::memcpy(hc, h.data() + nd + nf, sizeof(hc));
    r->result.assign(size_t(r->P.rows * r->scene->host.cols), 0.0);
    finalise_field(r->scene->host, r->P, h.data(), h.data() + nd, r->result.data());
    r->stats.bundles = int64_t(hc[C_BUNDLES]);
    r->stats.steps = int64_t(hc[C_STEPS]);
    r->stats.escapes = int64_t(hc[C_ESC]);
    r->stats.absorbed = int64_t(hc[C_ABS]);
    r->stats.truncated = int64_t(hc[C_TRUNC]);
    r->t_done = std::chrono::steady_clock::now();
    r->stats.wall_ms = std::chrono::duration<double, std::milli>(r->t_done - r->t_submit).count();
    r->state = 2;
    sv->used -= std::min(sv->used, r->bytes);
    if (r->stats.bundles != r->P.count) return set_error(MCG_EDEVICE, "request finished with bundles unaccounted for");
    return MCG_OK;
}

// Enqueue one precision group of a step on the server's stream.
template <typename F>
int enqueue_group(mcg_server* sv, StepSlot& slot, const std::vector<Request*>& reqs,
                  const std::vector<int64_t>& take, DevBuf<KParams<F>>& kpbuf,
                  DevBuf<unsigned long long>& nxbuf) {
    const int n = int(reqs.size());
    std::vector<mcg_scene*> sc(n);
    std::vector<HostProblem> P(n);
    for (int i = 0; i < n; ++i) {
        sc[i] = reqs[i]->scene.get();
        P[i] = reqs[i]->P;
    }
    BatchPlan B;
    int rc = plan_shape<F>(n, sc.data(), P.data(), &sv->tune, B);
    if (rc) return rc;
    int64_t handed = 0;
    for (int64_t t : take) handed += t;
    B.grid = int(std::max<int64_t>(1, std::min<int64_t>(B.grid, (handed + B.block - 1) / B.block)));
    if (nxbuf.n < size_t(n) && (rc = nxbuf.alloc(size_t(n) * 2))) return rc;
    std::vector<KParams<F>> K(n);
    for (int i = 0; i < n; ++i) {
        Request* r = reqs[i];
        K[i] = KParams<F>{};
        fill_params<F>(sc[i], r->P, K[i]);
        K[i].emit_cdf = r->cdf.p;
        K[i].direct = r->direct.p;
        K[i].diff = r->diff.p;
        K[i].counters = r->ctr.p;
        K[i].next = nxbuf.p + i;
        K[i].chunk = B.chunk;
        K[i].first = r->P.first + r->cursor;   // this step's share of the request's bundles
        K[i].count = take[i];
        K[i].lds_tally = B.tally ? 1 : 0;
        r->cursor += take[i];
        slot.handed.push_back({r, take[i]});
    }
    if (kpbuf.n < size_t(n) && (rc = kpbuf.alloc(size_t(n) * 2))) return rc;
    // synchronous copy: kpbuf of this slot is not in use on the device
    HIP_TRY(hipMemcpy(kpbuf.p, K.data(), K.size() * sizeof(KParams<F>), hipMemcpyHostToDevice));
    HIP_TRY(hipMemsetAsync(nxbuf.p, 0, size_t(n) * sizeof(unsigned long long), slot.stream));
    return launch_batch<F>(B, kpbuf.p, n, slot.stream);
}

// Hand out fair shares of the step budget and enqueue the step in `slot`;
// `other` is the slot whose step may still run.
int enqueue_step(mcg_server* sv, StepSlot& slot, StepSlot& other) {
    slot.handed.clear();
    slot.live = false;
    std::vector<Request*> g32, g64;
    std::vector<int64_t> t32, t64;
    int with_work = 0;
    for (Request* r : sv->active)
        if (r->cursor < r->P.count) ++with_work;
    if (!with_work && sv->fresh.empty()) return MCG_OK;
    if (with_work) {
        const int64_t share = std::max<int64_t>(1 << 16, sv->step_bundles / with_work);
        for (Request* r : sv->active) {
            const int64_t take = std::min(share, r->P.count - r->cursor);
            if (take <= 0) continue;
            (r->P.precision == MCG_FP64 ? g64 : g32).push_back(r);
            (r->P.precision == MCG_FP64 ? t64 : t32).push_back(take);
        }
    }
    int rc;
    if (!slot.stream) HIP_TRY(hipStreamCreateWithFlags(&slot.stream, hipStreamNonBlocking));
    if (!slot.e0) HIP_TRY(hipEventCreate(&slot.e0));
    if (!slot.e1) HIP_TRY(hipEventCreate(&slot.e1));
    if (!slot.ez) HIP_TRY(hipEventCreate(&slot.ez));
    // the requests the other step cleared are cleared before this step adds
    // to them (the wait covers the clearing, not the other step's kernel)
    if (other.ez_set) HIP_TRY(hipStreamWaitEvent(slot.stream, other.ez, 0));
    HIP_TRY(hipEventRecord(slot.e0, slot.stream));
    // requests admitted since the last step start from cleared tallies (no
    // earlier step touches them); the step's end event covers the clearing
    slot.ez_set = !sv->fresh.empty();
    for (Request* r : sv->fresh) {
        HIP_TRY(hipMemsetAsync(r->direct.p, 0, r->direct.n * sizeof(double), slot.stream));
        HIP_TRY(hipMemsetAsync(r->diff.p, 0, r->diff.n * sizeof(double), slot.stream));
        HIP_TRY(hipMemsetAsync(r->ctr.p, 0, C_COUNT * sizeof(unsigned long long), slot.stream));
    }
    sv->fresh.clear();
    if (slot.ez_set) HIP_TRY(hipEventRecord(slot.ez, slot.stream));
    if (!g32.empty() && (rc = enqueue_group<float>(sv, slot, g32, t32, slot.kp32, slot.nx32))) return rc;
    if (!g64.empty() && (rc = enqueue_group<double>(sv, slot, g64, t64, slot.kp64, slot.nx64))) return rc;
    HIP_TRY(hipEventRecord(slot.e1, slot.stream));
    slot.live = true;
    return MCG_OK;
}

// Wait for the step in `slot` and retire the requests it completed.
int finish_step(mcg_server* sv, StepSlot& slot) {
    if (!slot.live) return MCG_OK;
    HIP_TRY(hipEventSynchronize(slot.e1));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, slot.e0, slot.e1));
    sv->kernel_ms_total += ms;
    slot.live = false;
    for (auto& h : slot.handed) h.first->done += h.second;
    slot.handed.clear();
    std::vector<Request*> still;
    for (Request* r : sv->active) {
        if (r->done >= r->P.count) {
            int rc = retire(sv, r);
            if (rc) return rc;
            sv->steps_total += r->stats.steps;
            sv->completed.push_back(r->id);
        } else {
            still.push_back(r);
        }
    }
    sv->active.swap(still);
    return MCG_OK;
}

}  // namespace
}  // namespace mcg

// ====================================================================== C ABI
extern "C" {

int mcg_abi_version(void) { return MCG_ABI_VERSION; }

const char* mcg_last_error(void) { return mcg::t_err.c_str(); }

void mcg_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    // host evaluation of the device generator's rounds (for tests)
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3], k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = uint64_t(0xD2511F53u) * c0, p1 = uint64_t(0xCD9E8D57u) * c2;
        const uint32_t n0 = uint32_t(p1 >> 32) ^ c1 ^ k0, n2 = uint32_t(p0 >> 32) ^ c3 ^ k1;
        c0 = n0;
        c1 = uint32_t(p1);
        c2 = n2;
        c3 = uint32_t(p0);
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

int mcg_alias_build(const double* pdf, int64_t n, uint32_t* thresh, uint32_t* alias) {
    if (!pdf || !thresh || !alias) return mcg::set_error(MCG_EINVAL, "NULL argument");
    std::vector<uint64_t> t;
    int rc = mcg::build_alias(pdf, n, t);
    if (rc) return rc;
    for (int64_t i = 0; i < n; ++i) {
        thresh[i] = uint32_t(t[i]);
        alias[i] = uint32_t(t[i] >> 32);
    }
    return MCG_OK;
}

int mcg_material_compute(const mcg_material_spec* mat, mcg_material_tables* out) {
    if (!out) return mcg::set_error(MCG_EINVAL, "NULL output");
    mcg::HostScene s;
    int rc = mcg::build_material(mat, s);
    if (rc) return rc;
    const size_t n = size_t(s.nw * s.np) * sizeof(double);
    if (out->tau) std::memcpy(out->tau, s.tau.data(), n);
    if (out->vel) std::memcpy(out->vel, s.vel.data(), n);
    if (out->energy_pdf) std::memcpy(out->energy_pdf, s.energy.data(), n);
    if (out->flux_pdf) std::memcpy(out->flux_pdf, s.flux.data(), n);
    if (out->scat_pdf) std::memcpy(out->scat_pdf, s.scat.data(), n);
    out->energy_sum = s.energy_sum;
    out->flux_sum = s.flux_sum;
    out->scat_sum = s.scat_sum;
    out->cond = s.cond;
    out->nmodes = int64_t(s.mode_cell.size());
    return MCG_OK;
}

int mcg_problem_plan(const mcg_material_spec* mat, const mcg_domain_spec* dom,
                     const mcg_problem_spec* prob, mcg_problem_info* info, double* emit_weight,
                     int64_t* emit_count, int64_t nemitter_cap) {
    if (!info) return mcg::set_error(MCG_EINVAL, "NULL output");
    mcg::HostScene s;
    mcg::HostProblem p;
    int rc = mcg::build_material(mat, s);
    if (!rc) rc = mcg::build_domain(dom, s);
    if (!rc) rc = mcg::build_problem(prob, s, p);
    if (rc) return rc;
    info->nemit = p.nemit;
    info->rows = p.rows;
    info->cols = s.cols;
    info->maxloop = p.maxloop;
    info->nemitter = int64_t(s.emitters.size());
    info->power = p.power;
    info->weight_sum = p.weight_sum;
    for (int64_t i = 0; i < info->nemitter && i < nemitter_cap; ++i) {
        if (emit_weight) emit_weight[i] = s.emit_weight[i];
        if (emit_count) emit_count[i] = p.emit_count[i];
    }
    return MCG_OK;
}

int mcg_cell_volumes(const mcg_domain_spec* dom, double* out, int64_t cap, int64_t* ncols) {
    if (!ncols || (cap > 0 && !out)) return mcg::set_error(MCG_EINVAL, "NULL output");
    mcg::HostScene s;
    int rc = mcg::build_domain(dom, s);
    if (rc) return rc;
    *ncols = s.cols;
    for (int64_t i = 0; i < s.cols && i < cap; ++i) out[i] = s.cell_vol[i];
    return MCG_OK;
}

int mcg_scene_create(const mcg_material_spec* mat, const mcg_domain_spec* dom, mcg_scene** out) {
    if (!out) return mcg::set_error(MCG_EINVAL, "NULL output");
    *out = nullptr;
    std::unique_ptr<mcg_scene> sc(new (std::nothrow) mcg_scene());
    if (!sc) return mcg::set_error(MCG_ENOMEM, "out of host memory");
    int rc = mcg::build_material(mat, sc->host);
    if (!rc) rc = mcg::build_domain(dom, sc->host);
    if (!rc) rc = mcg::scene_upload(sc.get());
    if (rc) return rc;
    *out = sc.release();
    return MCG_OK;
}

int mcg_scene_plan(const mcg_scene* scene, const mcg_problem_spec* prob, mcg_problem_info* info) {
    if (!scene || !prob || !info) return mcg::set_error(MCG_EINVAL, "NULL argument");
    mcg::HostProblem p;
    int rc = mcg::build_problem(prob, scene->host, p);
    if (rc) return rc;
    info->nemit = p.nemit;
    info->rows = p.rows;
    info->cols = scene->host.cols;
    info->maxloop = p.maxloop;
    info->nemitter = int64_t(scene->host.emitters.size());
    info->power = p.power;
    info->weight_sum = p.weight_sum;
    return MCG_OK;
}

int mcg_scene_tune(mcg_scene* scene, const mcg_tune* tune) {
    if (!scene) return mcg::set_error(MCG_EINVAL, "NULL scene");
    mcg_tune t = tune ? *tune : mcg_tune{};
    if (t.block != 0 && (t.block < 64 || t.block > 1024 || t.block % 64 != 0))
        return mcg::set_error(MCG_EINVAL, "block must be a multiple of 64 from 64 to 1024");
    if (t.chunk < 0 || t.blocks_per_cu < 0 || t.per_thread < 0 || t.compact_every < 0 || t.live < 0)
        return mcg::set_error(MCG_EINVAL, "negative tuning value");
    if (t.per_thread != 0 && t.per_thread != 1 && t.per_thread != 2 && t.per_thread != 4 && t.per_thread != 8)
        return mcg::set_error(MCG_EINVAL, "per_thread must be 1, 2, 4 or 8");
    if (t.engine != MCG_ENGINE_HISTORY && t.engine != MCG_ENGINE_EVENT)
        return mcg::set_error(MCG_EINVAL, "unknown engine");
    if (t.chunk != 0 && t.chunk < 64) return mcg::set_error(MCG_EINVAL, "chunk must be at least 64");
    if (t.defer_walks < -1 || t.defer_walks > 1) return mcg::set_error(MCG_EINVAL, "defer_walks must be -1, 0 or 1");
    scene->tune = t;
    return MCG_OK;
}

int mcg_scene_destroy(mcg_scene* scene) {
    delete scene;
    return MCG_OK;
}

int mcg_solve(mcg_scene* scene, const mcg_problem_spec* prob, double* out, int64_t out_len,
              mcg_counts* counts, mcg_stats* stats) {
    if (!scene || !prob || !out) return mcg::set_error(MCG_EINVAL, "NULL argument");
    mcg::HostProblem p;
    int rc = mcg::build_problem(prob, scene->host, p);
    if (rc) return rc;
    if (out_len < p.rows * scene->host.cols) return mcg::set_error(MCG_ERANGE, "output buffer too small");
    if (p.count == 0) {
        std::vector<double> z(p.rows * scene->host.cols, 0.0), zf(p.rows * (scene->host.cols + 1), 0.0);
        mcg::finalise_field(scene->host, p, z.data(), zf.data(), out);
        if (stats) *stats = mcg_stats{};
        return MCG_OK;
    }
    if (scene->tune.engine == MCG_ENGINE_EVENT)
        return p.precision == MCG_FP64 ? mcg::run_solve_event<double>(scene, p, out, counts, stats)
                                       : mcg::run_solve_event<float>(scene, p, out, counts, stats);
    if (p.precision == MCG_FP64) return mcg::run_solve<double>(scene, p, out, counts, stats);
    return mcg::run_solve<float>(scene, p, out, counts, stats);
}

int mcg_solve_static(mcg_scene* scene, const mcg_problem_spec* prob, double* out, int64_t out_len,
                     mcg_stats* stats) {
    if (!scene || !prob || !out) return mcg::set_error(MCG_EINVAL, "NULL argument");
    mcg::HostProblem p;
    int rc = mcg::build_problem(prob, scene->host, p);
    if (rc) return rc;
    if (out_len < p.rows * scene->host.cols) return mcg::set_error(MCG_ERANGE, "output buffer too small");
    if (p.count == 0) {
        for (int64_t i = 0; i < p.rows * scene->host.cols; ++i) out[i] = 0.0;
        mcg::normalise_field(scene->host, p, out);
        if (stats) *stats = mcg_stats{};
        return MCG_OK;
    }
    if (p.precision == MCG_FP64) return mcg::run_solve_static<double>(scene, p, out, stats);
    return mcg::run_solve_static<float>(scene, p, out, stats);
}

int mcg_trace(mcg_scene* scene, const mcg_problem_spec* prob, int64_t bundle, double* rec,
              int64_t cap, int64_t* nrec) {
    if (!scene || !prob || !rec || !nrec || cap <= 0) return mcg::set_error(MCG_EINVAL, "NULL argument");
    mcg_problem_spec one = *prob;
    mcg::HostProblem p;
    int rc = mcg::build_problem(&one, scene->host, p);
    if (rc) return rc;
    if (bundle < 0 || bundle >= p.nemit) return mcg::set_error(MCG_ERANGE, "bundle outside [0, nemit)");
    one.first_bundle = bundle;
    one.bundle_count = 1;
    if ((rc = mcg::build_problem(&one, scene->host, p))) return rc;
    std::vector<double> out(size_t(p.rows * scene->host.cols));
    uint32_t c_nscat, c_nbdry, c_nstep, c_fate;
    mcg_counts counts{&c_nscat, &c_nbdry, &c_nstep, &c_fate};
    rc = p.precision == MCG_FP64
             ? mcg::run_solve<double>(scene, p, out.data(), &counts, nullptr, bundle, rec, cap)
             : mcg::run_solve<float>(scene, p, out.data(), &counts, nullptr, bundle, rec, cap);
    if (rc) return rc;
    *nrec = std::min<int64_t>(int64_t(c_nstep), cap);
    return MCG_OK;
}

int mcg_trace_from(mcg_scene* scene, const mcg_problem_spec* prob, int64_t prop_cell,
                   const double pos[3], const double* dir, double* rec, int64_t cap, int64_t* nrec) {
    if (!scene || !prob || !pos || !rec || !nrec || cap <= 0) return mcg::set_error(MCG_EINVAL, "NULL argument");
    const mcg::HostScene& h = scene->host;
    if (prob->first_bundle < 0) return mcg::set_error(MCG_ERANGE, "bundle id must be non-negative");
    mcg_problem_spec one = *prob;
    one.first_bundle = 0;
    one.bundle_count = 1;
    mcg::HostProblem p;
    int rc = mcg::build_problem(&one, h, p, true);
    if (rc) return rc;
    p.first = prob->first_bundle;   // the stream's id; it need not be an emitted bundle
    int mode = -1;
    if (prop_cell >= 0) {
        for (size_t k = 0; k < h.mode_cell.size(); ++k)
            if (h.mode_cell[k] == prop_cell) mode = int(k);
        if (mode < 0) return mcg::set_error(MCG_EINVAL, "prop has no energy pdf");
    }
    // Domain::locate: the first subdomain whose faces all hold pos within eps
    int sdom = -1;
    for (size_t si = 0; si < h.sdoms.size() && sdom < 0; ++si) {
        const mcg::DevSdom<double>& S = h.sdoms[si];
        bool in = true;
        for (int b = S.pbeg; b < S.pend && in; ++b) {
            const mcg::DevPlane<double>& P = h.planes[b];
            in = P.n[0] * pos[0] + P.n[1] * pos[1] + P.n[2] * pos[2] + P.c >= -S.eps;
        }
        if (in && S.pore >= 0) {
            const mcg::DevPore<double>& Q = h.pores[S.pore];
            const double dx = pos[0] - Q.cx, dy = pos[1] - Q.cy, rr = Q.r - Q.eps;
            in = !(dx * dx + dy * dy < rr * rr);
        }
        if (in) sdom = int(si);
    }
    if (sdom < 0) return mcg::set_error(MCG_EINVAL, "Position not inside domain");
    if (dir && !(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2] > 0.0))
        return mcg::set_error(MCG_EINVAL, "direction must be nonzero");
    if (p.precision == MCG_FP64) {
        mcg::TraceStart<double> s{};
        for (int k = 0; k < 3; ++k) { s.pos[k] = pos[k]; s.dir[k] = dir ? dir[k] : 0.0; }
        s.has_dir = dir ? 1 : 0;
        s.mode = mode;
        s.sdom = sdom;
        return mcg::run_trace_from<double>(scene, p, s, rec, cap, nrec);
    }
    mcg::TraceStart<float> s{};
    for (int k = 0; k < 3; ++k) { s.pos[k] = pos[k]; s.dir[k] = dir ? dir[k] : 0.0; }
    s.has_dir = dir ? 1 : 0;
    s.mode = mode;
    s.sdom = sdom;
    return mcg::run_trace_from<float>(scene, p, s, rec, cap, nrec);
}

int mcg_batch_solve(int32_t n, mcg_scene* const* scenes, const mcg_problem_spec* probs,
                    double* const* outs, const int64_t* out_lens, mcg_stats* stats,
                    const mcg_tune* tune) {
    if (n < 0 || (n > 0 && (!scenes || !probs || !outs || !out_lens)))
        return mcg::set_error(MCG_EINVAL, "NULL argument");
    if (n == 0) return MCG_OK;
    std::vector<mcg::HostProblem> P(n);
    std::vector<int> by_prec[2];
    for (int32_t i = 0; i < n; ++i) {
        if (!scenes[i] || !outs[i]) return mcg::set_error(MCG_EINVAL, "NULL scene or output");
        int rc = mcg::build_problem(&probs[i], scenes[i]->host, P[i]);
        if (rc) return rc;
        if (out_lens[i] < P[i].rows * scenes[i]->host.cols)
            return mcg::set_error(MCG_ERANGE, "output buffer too small");
        by_prec[P[i].precision == MCG_FP64 ? 1 : 0].push_back(i);
    }
    // one launch per precision present
    for (int g = 0; g < 2; ++g) {
        const std::vector<int>& idx = by_prec[g];
        if (idx.empty()) continue;
        std::vector<mcg_scene*> sc;
        std::vector<mcg::HostProblem> pp;
        std::vector<double*> oo;
        std::vector<mcg_stats> ss(idx.size());
        for (int i : idx) {
            sc.push_back(scenes[i]);
            pp.push_back(P[i]);
            oo.push_back(outs[i]);
        }
        int rc = g == 1 ? mcg::run_batch<double>(int(idx.size()), sc.data(), pp.data(), oo.data(), ss.data(), tune)
                        : mcg::run_batch<float>(int(idx.size()), sc.data(), pp.data(), oo.data(), ss.data(), tune);
        if (rc) return rc;
        if (stats)
            for (size_t k = 0; k < idx.size(); ++k) stats[idx[k]] = ss[k];
    }
    return MCG_OK;
}

struct mcg_batch {
    std::unique_ptr<mcg::GraphBatchBase> impl;
};

int mcg_batch_create(int32_t n, mcg_scene* const* scenes, const mcg_problem_spec* probs,
                     const mcg_tune* tune, mcg_batch** out) {
    if (!out || n <= 0 || !scenes || !probs) return mcg::set_error(MCG_EINVAL, "NULL argument or empty batch");
    *out = nullptr;
    std::vector<mcg::HostProblem> P(n);
    for (int32_t i = 0; i < n; ++i) {
        if (!scenes[i]) return mcg::set_error(MCG_EINVAL, "NULL scene");
        if (scenes[i]->device != scenes[0]->device) return mcg::set_error(MCG_EINVAL, "batch scenes on different devices");
        int rc = mcg::build_problem(&probs[i], scenes[i]->host, P[i]);
        if (rc) return rc;
        if (P[i].precision != P[0].precision) return mcg::set_error(MCG_EINVAL, "a graph batch takes one precision");
    }
    std::unique_ptr<mcg_batch> b(new (std::nothrow) mcg_batch());
    if (!b) return mcg::set_error(MCG_ENOMEM, "out of host memory");
    int rc;
    if (P[0].precision == MCG_FP64) {
        auto* g = new (std::nothrow) mcg::GraphBatch<double>();
        if (!g) return mcg::set_error(MCG_ENOMEM, "out of host memory");
        b->impl.reset(g);
        g->n = n;
        g->sc.assign(scenes, scenes + n);
        g->P = P;
        rc = g->create(tune);
    } else {
        auto* g = new (std::nothrow) mcg::GraphBatch<float>();
        if (!g) return mcg::set_error(MCG_ENOMEM, "out of host memory");
        b->impl.reset(g);
        g->n = n;
        g->sc.assign(scenes, scenes + n);
        g->P = P;
        rc = g->create(tune);
    }
    if (rc) return rc;
    *out = b.release();
    return MCG_OK;
}

int mcg_batch_run(mcg_batch* batch, const uint64_t* seeds, double* const* outs, const int64_t* out_lens,
                  mcg_stats* stats) {
    if (!batch || !batch->impl || !outs || !out_lens) return mcg::set_error(MCG_EINVAL, "NULL argument");
    for (int i = 0; i < batch->impl->n; ++i) {
        if (!outs[i]) return mcg::set_error(MCG_EINVAL, "NULL output");
        if (out_lens[i] < batch->impl->need(i)) return mcg::set_error(MCG_ERANGE, "output buffer too small");
    }
    return batch->impl->run(seeds, outs, stats);
}

int mcg_batch_destroy(mcg_batch* batch) {
    delete batch;
    return MCG_OK;
}

int mcg_autotune(mcg_scene* scene, const mcg_problem_spec* prob, int64_t sample_bundles,
                 mcg_tune* best, double* best_steps_per_s) {
    if (!scene || !prob) return mcg::set_error(MCG_EINVAL, "NULL argument");
    mcg::HostProblem p;
    int rc = mcg::build_problem(prob, scene->host, p);
    if (rc) return rc;
    const int64_t n = sample_bundles > 0 ? std::min(sample_bundles, p.count)
                                         : std::min<int64_t>(p.count, int64_t(1) << 20);
    p.count = n;
    std::vector<double> out(size_t(p.rows * scene->host.cols));
    const mcg_tune saved = scene->tune;
    mcg_tune top = saved;
    double top_rate = -1.0;
    // candidates: the history kernel's block, table placement and grab size;
    // the event engine's bundles in flight, bundles per thread and rounds
    // between regroupings
    std::vector<mcg_tune> cand;
    static const int blocks[] = {64, 128, 256, 512, 1024};
    static const int tables[] = {1, -1};
    static const int chunks[] = {64, 256, 1024};
    for (int b : blocks)
        for (int t : tables)
            for (int c : chunks) {
                mcg_tune x{};
                x.block = b;
                x.chunk = c;
                x.lds_tables = t;
                x.engine = MCG_ENGINE_HISTORY;
                cand.push_back(x);
            }
    static const int64_t lives[] = {int64_t(1) << 18, int64_t(1) << 20};
    static const int per_thread[] = {1, 4};
    static const int compact[] = {1, 4};
    for (int64_t lv : lives)
        for (int pt : per_thread)
            for (int ce : compact) {
                mcg_tune x{};
                x.engine = MCG_ENGINE_EVENT;
                x.live = std::min<int64_t>(lv, n);
                x.per_thread = pt;
                x.compact_every = ce;
                cand.push_back(x);
            }
    auto measure = [&](const mcg_tune& x) {
        scene->tune = x;
        mcg_stats st{};
        double ms = 1e300;
        for (int rep = 0; rep < 2; ++rep) {   // first run warms up the shape
            int r;
            if (x.engine == MCG_ENGINE_EVENT)
                r = p.precision == MCG_FP64 ? mcg::run_solve_event<double>(scene, p, out.data(), nullptr, &st)
                                            : mcg::run_solve_event<float>(scene, p, out.data(), nullptr, &st);
            else
                r = p.precision == MCG_FP64 ? mcg::run_solve<double>(scene, p, out.data(), nullptr, &st)
                                            : mcg::run_solve<float>(scene, p, out.data(), nullptr, &st);
            if (r) return r;
            ms = std::min(ms, st.kernel_ms);
        }
        const double rate = double(st.steps) / (ms * 1e-3);
        if (rate > top_rate) {
            top_rate = rate;
            top = x;
        }
        return int(MCG_OK);
    };
    for (const mcg_tune& x : cand)
        if ((rc = measure(x))) break;
    // a scene with 2-D / 3-D grids: the best history shape with its tally
    // walks deferred and walked inside the step
    if (!rc && scene->host.has_grid && top.engine == MCG_ENGINE_HISTORY) {
        const mcg_tune shape = top;
        for (int d : {1, -1}) {
            mcg_tune x = shape;
            x.defer_walks = d;
            if ((rc = measure(x))) break;
        }
    }
    if (rc) {
        scene->tune = saved;
        return rc;
    }
    scene->tune = top;
    if (best) *best = top;
    if (best_steps_per_s) *best_steps_per_s = top_rate;
    return MCG_OK;
}

int mcg_autotune_batch(int32_t n, mcg_scene* const* scenes, const mcg_problem_spec* probs,
                       double sample_fraction, mcg_tune* best, double* best_steps_per_s) {
    if (n <= 0 || !scenes || !probs) return mcg::set_error(MCG_EINVAL, "NULL argument");
    const double frac = sample_fraction > 0.0 ? std::min(sample_fraction, 1.0) : 0.1;
    std::vector<mcg::HostProblem> P(n);
    std::vector<mcg_scene*> sc(scenes, scenes + n);
    std::vector<std::vector<double>> outs(n);
    std::vector<double*> optr(n);
    int prec = -1;
    for (int i = 0; i < n; ++i) {
        int rc = mcg::build_problem(&probs[i], scenes[i]->host, P[i]);
        if (rc) return rc;
        if (prec >= 0 && P[i].precision != prec)
            return mcg::set_error(MCG_EINVAL, "autotune a batch of one precision at a time");
        prec = P[i].precision;
        P[i].count = std::max<int64_t>(std::min<int64_t>(P[i].count, 256), int64_t(double(P[i].count) * frac));
        outs[i].resize(size_t(P[i].rows * scenes[i]->host.cols));
        optr[i] = outs[i].data();
    }
    std::vector<mcg_stats> st(n);
    mcg_tune top{};
    double top_rate = -1.0;
    auto measure = [&](const mcg_tune& cand) {
        double ms = 1e300;
        int64_t steps = 0;
        for (int rep = 0; rep < 2; ++rep) {
            int rc = prec == MCG_FP64
                         ? mcg::run_batch<double>(n, sc.data(), P.data(), optr.data(), st.data(), &cand)
                         : mcg::run_batch<float>(n, sc.data(), P.data(), optr.data(), st.data(), &cand);
            if (rc) return rc;
            ms = std::min(ms, st[0].kernel_ms);
            steps = 0;
            for (const mcg_stats& x : st) steps += x.steps;
        }
        const double rate = double(steps) / (ms * 1e-3);
        if (rate > top_rate) {
            top_rate = rate;
            top = cand;
        }
        return int(MCG_OK);
    };
    static const int blocks[] = {128, 256, 512, 1024};
    static const int tables[] = {1, -1};
    static const int chunks[] = {64, 256, 1024};
    int rc;
    for (int b : blocks)
        for (int t : tables)
            for (int c : chunks) {
                mcg_tune cand{};
                cand.block = b;
                cand.chunk = c;
                cand.lds_tables = t;
                if ((rc = measure(cand))) return rc;
            }
    // FP32 batches with 2-D / 3-D grids: the best shape with its tally walks
    // deferred and walked inside the step (FP64 batches walk inside the step)
    bool grid = false;
    for (int i = 0; i < n; ++i) grid = grid || scenes[i]->host.has_grid;
    if (grid && prec != MCG_FP64) {
        const mcg_tune shape = top;
        for (int d : {1, -1}) {
            mcg_tune cand = shape;
            cand.defer_walks = d;
            if ((rc = measure(cand))) return rc;
        }
    }
    if (best) *best = top;
    if (best_steps_per_s) *best_steps_per_s = top_rate;
    return MCG_OK;
}


int mcg_server_create(uint64_t hbm_budget_bytes, int64_t step_bundles, const mcg_tune* tune,
                      mcg_server** out) {
    if (!out) return mcg::set_error(MCG_EINVAL, "NULL output");
    *out = nullptr;
    std::unique_ptr<mcg_server> sv(new (std::nothrow) mcg_server());
    if (!sv) return mcg::set_error(MCG_ENOMEM, "out of host memory");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return mcg::set_error(MCG_EDEVICE, "no HIP device available");
    sv->budget = hbm_budget_bytes > 0 ? size_t(hbm_budget_bytes) : size_t(200) << 30;
    sv->step_bundles = step_bundles > 0 ? step_bundles : (int64_t(1) << 24);
    if (tune) sv->tune = *tune;
    HIP_TRY(hipGetDevice(&sv->device));
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, sv->device));
    sv->num_cu = prop.multiProcessorCount;
    sv->lds_limit = prop.sharedMemPerBlock;
    for (mcg::StepSlot& sl : sv->slot) HIP_TRY(hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking));
    const unsigned hc = std::thread::hardware_concurrency();
    const int nworkers = int(std::max(1u, std::min(8u, hc > 1 ? hc / 2 : 1u)));
    try {
        for (int i = 0; i < nworkers; ++i) sv->workers.emplace_back(mcg::build_worker, sv.get());
    } catch (const std::exception&) {
        mcg::stop_workers(sv.get());
        return mcg::set_error(MCG_ENOMEM, "cannot start the server's build threads");
    }
    *out = sv.release();
    return MCG_OK;
}

int mcg_server_submit(mcg_server* sv, const mcg_material_spec* mat, const mcg_domain_spec* dom,
                      const mcg_problem_spec* prob, int64_t* request_id) {
    if (!sv || !request_id) return mcg::set_error(MCG_EINVAL, "NULL argument");
    std::unique_ptr<mcg::Request> r(new (std::nothrow) mcg::Request());
    if (!r) return mcg::set_error(MCG_ENOMEM, "out of host memory");
    r->scene.reset(new (std::nothrow) mcg_scene());
    if (!r->scene) return mcg::set_error(MCG_ENOMEM, "out of host memory");
    if (!mat || !prob) return mcg::set_error(MCG_EINVAL, "NULL argument");
    // what the copies below rely on; the rest of the material is checked by its build
    if (!mat->disp || !mat->relax) return mcg::set_error(MCG_EINVAL, "material spec is NULL");
    if (mat->nw <= 0 || mat->np <= 0) return mcg::set_error(MCG_EINVAL, "Invalid dispersion file");
    // the domain is built now (its spec's arrays are the caller's); the
    // material, problem and device image on a build thread, from copies
    int rc = mcg::build_domain(dom, r->scene->host);
    if (rc) return rc;
    const mcg::HostScene& h = r->scene->host;
    {
        // a lower bound of the request's device bytes: geometry and tallies
        const int64_t rows = prob->tally == MCG_TEMP ? 1 : prob->tally == MCG_FLUX ? 3 : prob->tally == MCG_MULTI ? 4
                             : prob->tally == MCG_CUMTEMP ? std::max<int64_t>(1, prob->size + 1)
                                                          : 3 * std::max<int64_t>(1, prob->size + 1);
        const size_t lower = h.planes.size() * (sizeof(mcg::DevPlane<double>) + sizeof(mcg::DevPlane<float>)) +
                             h.sdoms.size() * (sizeof(mcg::DevSdom<double>) + sizeof(mcg::DevSdom<float>)) +
                             sizeof(double) * size_t(rows) * (2 * size_t(h.cols) + 1);
        if (lower > sv->budget) return mcg::set_error(MCG_ENOMEM, "request exceeds the device memory budget");
    }
    const size_t nd = size_t(mat->nw) * size_t(2 + 2 * mat->np), nr = size_t(mat->np) * 8;
    r->disp.assign(mat->disp, mat->disp + nd);
    r->relax.assign(mat->relax, mat->relax + nr);
    r->mspec = *mat;
    r->mspec.disp = r->disp.data();
    r->mspec.relax = r->relax.data();
    r->pspec = *prob;
    r->id = sv->next_id++;
    r->t_submit = std::chrono::steady_clock::now();
    *request_id = r->id;
    mcg::Request* raw = r.get();
    sv->queue.push_back(raw);
    sv->all[r->id] = std::move(r);
    {
        std::lock_guard<std::mutex> lk(sv->mu);
        sv->todo.push_back(raw);
    }
    sv->cv_todo.notify_one();
    return MCG_OK;
}

int mcg_server_step(mcg_server* sv, int32_t* active, int32_t* queued, int32_t* completed) {
    if (!sv) return mcg::set_error(MCG_EINVAL, "NULL server");
    int rc;
    sv->completed.clear();
    // admission in arrival order while the budget allows (a request's build
    // is waited for when it reaches the head of the queue; one that failed,
    // or that could never fit, is completed with its error)
    while (!sv->queue.empty()) {
        mcg::Request* r = sv->queue.front();
        mcg::wait_built(sv, r);
        if (r->built > 0 && r->bytes > sv->budget) {
            r->built = -1;
            r->err_code = MCG_ENOMEM;
            r->err_msg = "request exceeds the device memory budget";
        }
        if (r->built < 0) {
            sv->queue.pop_front();
            r->state = 3;
            sv->completed.push_back(r->id);
            continue;
        }
        if (sv->used + r->bytes > sv->budget) break;
        sv->queue.pop_front();
        if ((rc = mcg::admit(sv, r))) {
            // the request fails with the error (its block, if any, is freed
            // once it is polled or cancelled); the others go on
            r->state = 3;
            r->err_code = rc;
            r->err_msg = mcg::t_err;
            sv->completed.push_back(r->id);
        }
    }
    // enqueue the next step behind the one in flight, then finish that one
    const int nxt = 1 - sv->cur;
    if ((rc = mcg::enqueue_step(sv, sv->slot[nxt], sv->slot[sv->cur]))) return rc;
    if ((rc = mcg::finish_step(sv, sv->slot[sv->cur]))) return rc;
    sv->cur = nxt;
    if (!sv->slot[sv->cur].live) {
        // nothing in flight: retire what the step just enqueued would have, and
        // release the memory of retired requests
        bool again = false;
        for (mcg::Request* r : sv->active) again = again || r->cursor < r->P.count;
        if (!again && !sv->graveyard.empty()) mcg::bury(sv);
    } else if (sv->graveyard.size() > 4096) {
        mcg::bury(sv);   // bound host-side bookkeeping under continuous load
    }
    if (active) *active = int32_t(sv->active.size());
    if (queued) *queued = int32_t(sv->queue.size());
    if (completed) *completed = int32_t(sv->completed.size());
    return MCG_OK;
}

int mcg_server_completed(const mcg_server* sv, int64_t* ids, int64_t cap, int64_t* n) {
    if (!sv || !n) return mcg::set_error(MCG_EINVAL, "NULL argument");
    *n = int64_t(sv->completed.size());
    for (int64_t i = 0; i < *n && i < cap; ++i) ids[i] = sv->completed[size_t(i)];
    return *n > cap ? mcg::set_error(MCG_ERANGE, "id buffer too small") : MCG_OK;
}

int mcg_server_poll(mcg_server* sv, int64_t id, double* out, int64_t out_len, mcg_stats* stats) {
    if (!sv) return mcg::set_error(MCG_EINVAL, "NULL server");
    auto it = sv->all.find(id);
    if (it == sv->all.end()) return mcg::set_error(MCG_ENOTFOUND, "unknown request id");
    mcg::Request* r = it->second.get();
    if (r->state == 3) {
        const int code = r->err_code;
        const std::string msg = r->err_msg;
        sv->graveyard.push_back(std::move(it->second));
        sv->all.erase(it);
        return mcg::set_error(code, msg);
    }
    if (r->state != 2) return MCG_EAGAIN;
    if (out) {
        if (out_len < int64_t(r->result.size())) return mcg::set_error(MCG_ERANGE, "output buffer too small");
        std::memcpy(out, r->result.data(), r->result.size() * sizeof(double));
    }
    if (stats) *stats = r->stats;
    sv->graveyard.push_back(std::move(it->second));
    sv->all.erase(it);
    return MCG_OK;
}

int mcg_server_cancel(mcg_server* sv, int64_t id) {
    if (!sv) return mcg::set_error(MCG_EINVAL, "NULL server");
    auto it = sv->all.find(id);
    if (it == sv->all.end()) return mcg::set_error(MCG_ENOTFOUND, "unknown request id");
    mcg::Request* r = it->second.get();
    if (r->state == 0) {
        // a build thread may hold it: wait for the build before letting go
        mcg::wait_built(sv, r);
        sv->queue.erase(std::remove(sv->queue.begin(), sv->queue.end(), r), sv->queue.end());
    } else if (r->state == 1) {
        // the request may still have work in flight: wait for it first
        int rc = mcg::finish_step(sv, sv->slot[sv->cur]);
        if (rc) return rc;
        if (r->state == 1) {
            sv->active.erase(std::remove(sv->active.begin(), sv->active.end(), r), sv->active.end());
            sv->fresh.erase(std::remove(sv->fresh.begin(), sv->fresh.end(), r), sv->fresh.end());
            sv->used -= std::min(sv->used, r->bytes);
        }
    }
    sv->graveyard.push_back(std::move(it->second));
    sv->all.erase(it);
    return MCG_OK;
}

int mcg_server_info(const mcg_server* sv, mcg_server_stats* out) {
    if (!sv || !out) return mcg::set_error(MCG_EINVAL, "NULL argument");
    out->queued = int64_t(sv->queue.size());
    out->active = int64_t(sv->active.size());
    out->bytes_used = uint64_t(sv->used);
    out->bytes_budget = uint64_t(sv->budget);
    out->kernel_ms = sv->kernel_ms_total;
    out->steps = sv->steps_total;
    return MCG_OK;
}

int mcg_server_destroy(mcg_server* sv) {
    if (!sv) return MCG_OK;
    mcg::stop_workers(sv);   // builds in progress finish first
    for (mcg::StepSlot& sl : sv->slot)
        if (sl.stream) (void)hipStreamSynchronize(sl.stream);
    for (auto& kv : sv->all) sv->graveyard.push_back(std::move(kv.second));
    sv->all.clear();
    mcg::bury(sv);
    delete sv;
    return MCG_OK;
}

}  // extern "C"
