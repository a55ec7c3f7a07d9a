// spss_kernels_fast.hip -- FP32 cell-list path (SPSS_PRECISION_FP32C), models 3 and 4: host side.
//
// Data path of one RHS evaluation (stage state in FP64 AoS, original order):
//   fk_keys      cell of every particle on a grid of cells 1/(k*xsplit*nc) wide and 1/(k*nc) high,
//                histogram by atomics
//   scan         exclusive scan of the cell counts (two kernels)
//   fk_scatter   unstable bucket fill
//   fk_rank      stable position inside the cell (ascending particle index) -> sorted order,
//                and the snapshot: FP32 offsets of the particle inside its cell and its
//                unit phasor (cos phi, sin phi); omega for the SRK force
//   fk_pairs     one workgroup per tile of W cells of one row: the (2ky+1) rows x (W+2kx)
//                cells around it (kx = k*xsplit, ky = k) are staged into LDS with x re-based to
//                the tile (FP64 rounding, once per staged particle); each lane walks the (2ky+1) contiguous
//                row ranges of its particle in FP32 and accumulates the phasors of its
//                neighbours in FP64 (coupling = Im(e^{-i(phi+alpha)} sum e^{i phi_j})).
//                Pairs whose FP32 distance lies in a
//                guard band around rho^2 are re-tested with the reference's FP64 ClassC
//                formula, so the neighbour sets equal the FP64 ones exactly.
//   self terms (v0 cos phi, v0 sin phi, omega, -xi omega + sigma A / n) in FP64.
//
// Fused stochastic steps keep the state resident in cell order (see spss_fast_kernels.hpp):
//   fr_pairs (MODE 3 phasor sums; its first workgroups scan the arrival counts the previous step
//   left, with a decoupled look-back, into the new cell starts) -> fr_far (the far arrivals' slots
//   and the tiles that overflowed the LDS stage; usually none) -> fr_scatter (FP64 Euler/SRK update
//   written to the new slots, the next step's moves), three launches per step and no host
//   synchronisation.
#include <algorithm>
#include <cmath>
#include <string>
#include <utility>

#include "../../include/spss_hip.h"
#include "spss_fast.hpp"
#include "spss_launch.hpp"

#include "spss_fast_kernels.hpp"

namespace spss {

namespace {
thread_local std::string g_fast_err;
}  // namespace

struct FastPath {
  int n = 0, s = 4;
  double rho = 0.0, rho2 = 0.0;
  // grid: ncx x ncy cells of Hx x Hy per replica, stencil (2ky+1) x (2kx+1) cells; k the subdivision,
  // xsplit the columns of a square cell
  int nc = 1, k = 1, xsplit = 1, ncx = 1, ncy = 1, kx = 1, ky = 1, cpr = 1, cells = 1, nsb = 1;
  double Hx = 1.0, Hy = 1.0;
  float r2c = 0.f, r2h = 0.f, r2k = 0.f;
  // replicas (one for a single system) and the tiles of the pair kernel
  int R = 1;
  std::vector<int> rep_first_h, W_h;
  int *rep_first = nullptr, *rep_chunk = nullptr;
  int4 *tiles = nullptr;
  int4 *ovf_list = nullptr;  // work split off overflowing tiles
  int *ovf_count = nullptr;
  int ntiles = 0, tiles_cap = 0;
  // single systems: per-step row tiles of the resident order (cmax slots per row of cells), and
  // the row tiles of the rebuilt cell list (RHS, pair export), so that every pair-kernel launch
  // of a single system stages the same kind of tile
  int4 *rtiles = nullptr, *btiles = nullptr;
  int rt_cmax = 0, rt_wcap = 0;
  Coef *coef = nullptr;                    // per-replica coefficients (batches)
  unsigned long long *step_dev = nullptr;  // device step counter (batches, graph-replayed steps)
  int gen = 0;                             // bumped whenever launch shapes change (tile table)
  // graphs of fused steps of a single system: sg[0][p] runs kGraphSteps steps, sg[1][p] one
  // step, from buffer parity p (fast_buffer_parity); all were captured with the key below
  struct StepGraph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
  } sg[2][4];
  int sg_srk = -1, sg_gen = -1, sg_higher = -1;
  double sg_dt = 0.0, sg_dphi = 0.0, sg_friction = 0.0;
  uint64_t sg_seed = 0;
  const double *sg_state = nullptr, *sg_nat = nullptr;
  double *sg_compl = nullptr;
  long sg_captures = 0;  // step graphs captured since creation
  double *zlast = nullptr, *racc = nullptr;
  // rebuilt cell list (RHS, pair export, entry into resident mode)
  int *key = nullptr, *slot = nullptr, *count = nullptr, *start = nullptr, *tmp = nullptr, *sorted = nullptr;
  int *blocksum = nullptr;
  float4 *snap = nullptr;
  float *omega_f = nullptr;
  // resident copy of the fused steps (double-buffered where a step permutes)
  bool resident = false;
  bool resident_nat = false;   // rnat holds natural frequencies
  bool romega_valid = false;   // romega holds the FP32 omega of the resident order (SRK)
  int cur = 0;
  double *rstate[2] = {nullptr, nullptr};
  int *rid[2] = {nullptr, nullptr};
  double *rnat[2] = {nullptr, nullptr};
  int *rstart[2] = {nullptr, nullptr};
  float4 *rsnap = nullptr;
  float *romega = nullptr;
  double2 *ssum = nullptr, *tsum = nullptr;
  int *scnt = nullptr;
  int *cnt9[2] = {nullptr, nullptr};  // move counts per (cell, direction), by step parity (cur)
  int *newcnt = nullptr, *farcnt = nullptr;
  // moves of the current step by slot, double-buffered (mvc): every fr_scatter computes the
  // next step's moves and counts from the new state; mv_ready says they are there, for the
  // dt, stepper and v0 recorded (another step size or stepper needs fr_move again)
  int *mvb[2] = {nullptr, nullptr};
  int mvc = 0;
  // the arrivals scan: look-back status words per scan workgroup, its finished count, the
  // far movers of the step (slots) and their number, a flag for a look-back wait that ran out
  unsigned long long *lb_status = nullptr;
  // far movers by step parity (cur): the step's list and count, the next step's filled by fr_scatter
  int *farlist[2] = {nullptr, nullptr}, *farlist_n[2] = {nullptr, nullptr};
  int *lb_error = nullptr;
  int *farslot = nullptr;  // each far mover's slot among its destination's far arrivals
  bool mv_ready = false;
  bool mv_srk = false;
  double mv_dt = 0.0, mv_v0 = 0.0;
  hipStream_t st = nullptr;
  KernelTimer *timer = nullptr;
};

// ---------------------------------------------------------------------------
// host side

const char *fast_last_error() { return g_fast_err.c_str(); }

static int ffail(const std::string &m) {
  g_fast_err = m;
  return SPSS_ERR_DEVICE;
}

#define FCHK(call)                                                                         \
  do {                                                                                     \
    hipError_t e_ = (call);                                                                \
    if (e_ != hipSuccess) return ffail(std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

template <class T>
static hipError_t falloc(T **p, size_t count) {
  return hipMalloc((void **)p, std::max<size_t>(count, 1) * sizeof(T));
}

// widest tile: its staged (2ky+1) x (W+2kx) cells fit the pair kernel's cell tables, and at most
// 48 * xsplit staged columns (48 square cells: about the width of a dense row's full tile)
static int max_tile_width(const FastPath *f) {
  return std::min(kMaxCells / (2 * f->ky + 1), 48 * f->xsplit) - 2 * f->kx;
}

// default tile width of a replica: about 0.8 particles per thread at its mean density, with
// the staged (2ky+1) x (W+2kx) cells and their row padding within 85% of the LDS stage
static int default_tile_width(const FastPath *f, double per_cell) {
  const int nrows = 2 * f->ky + 1;
  // about 80% of the workgroup's lanes: a tile just above kFB particles would need a second pass
  // in which one wave walks the whole stencil for a handful of particles
  int W = (int)(0.8 * kFB / std::max(per_cell, 1e-9));
  const int wstage = (int)((0.85 * kStageMax - nrows * kPad) / std::max(per_cell * nrows, 1e-9)) - 2 * f->kx;
  W = std::min(W, std::max(wstage, 1));
  return std::max(1, std::min(W, std::min(max_tile_width(f), f->ncx)));
}

// tile table for per-replica tile widths, and the guard band of the widest span any launch can
// stage: the FP32 distance error bound (x coordinates <= (W+2kx+1) cells from the staged origin,
// y <= 2ky+2 cells, rounded once; offsets < one cell) times a safety factor of 8.  A single
// system's resident steps stage per-step row tiles of up to rt_wcap columns, so their width
// counts too.
static int set_tiles(FastPath *f, const std::vector<int> &W) {
  std::vector<int4> t;
  int wmax = f->rtiles ? f->rt_wcap : 1;
  for (int r = 0; r < f->R; ++r) {
    const int w = std::max(1, std::min(W[r], std::min(max_tile_width(f), f->ncx)));
    wmax = std::max(wmax, w);
    for (int row = 0; row < f->ncy; ++row)
      for (int ca = 0; ca < f->ncx; ca += w) t.push_back(make_int4(r * f->cpr, row, ca, std::min(ca + w, f->ncx)));
  }
  if ((int)t.size() > f->tiles_cap) {
    hipFree(f->tiles);
    f->tiles = nullptr;
    FCHK(hipMalloc(&f->tiles, sizeof(int4) * t.size()));
    f->tiles_cap = (int)t.size();
  }
  FCHK(hipMemcpy(f->tiles, t.data(), sizeof(int4) * t.size(), hipMemcpyHostToDevice));
  f->ntiles = (int)t.size();
  f->W_h = W;
  ++f->gen;
  const double span = std::max((double)(wmax + 2 * f->kx + 1) * f->Hx, (double)(2 * f->ky + 2) * f->Hy);
  const double eps = std::ldexp(1.0, -24);
  const double dxerr = 2.0 * span * eps + 2.0 * std::max(f->Hx, f->Hy) * eps;
  double g = 8.0 * (2.0 * dxerr / f->rho + 4.0 * eps);
  g = std::max(g, 64.0 * eps);
  f->r2c = (float)f->rho2;
  f->r2h = (float)(f->rho2 * g);
  f->r2k = (float)std::ldexp(1.0, std::ilogb(4.0 / (double)f->r2h) + 1);  // power of two, r2k * r2h > 4
  return SPSS_OK;
}

int fast_create_replicas(FastPath **out, hipStream_t st, const DevParams &P, double rho, int nc, int subdiv,
                         int xsplit, int R, const int *rep_n, const Coef *coef) {
  *out = nullptr;
  if (subdiv < 1 || subdiv > 3) {
    g_fast_err = "cell_subdivision must be 1, 2 or 3";
    return SPSS_ERR_INVALID;
  }
  if (xsplit != 1 && xsplit != 2 && xsplit != 4) {
    g_fast_err = "the cell x split must be 1, 2 or 4";
    return SPSS_ERR_INVALID;
  }
  if (P.xs != 1.0 || P.ys != 1.0) {
    g_fast_err = "the FP32C cell list assumes the unit box of the reference's cell grid";
    return SPSS_ERR_UNSUPPORTED;
  }
  if (nc < 3) {
    g_fast_err = "the FP32C cell list needs int(1/rho) >= 3: on narrower grids the reference visits cells "
                 "more than once (use the FP64 precision there)";
    return SPSS_ERR_UNSUPPORTED;
  }
  const long cpr = (long)nc * subdiv * xsplit * nc * subdiv;
  if (R < 1 || cpr * R >= (1L << 27)) {
    g_fast_err = "the FP32C cell list supports at most 2^27 cells over all replicas";
    return SPSS_ERR_UNSUPPORTED;
  }
  FastPath *f = new FastPath();
  f->n = P.n;
  f->s = P.s;
  f->rho = rho;
  f->rho2 = rho * rho;
  f->nc = nc;
  f->k = subdiv;
  f->xsplit = xsplit;
  f->ncy = nc * subdiv;
  f->ncx = f->ncy * xsplit;
  f->ky = subdiv;
  f->kx = subdiv * xsplit;
  f->Hx = 1.0 / (double)f->ncx;
  f->Hy = 1.0 / (double)f->ncy;
  f->cpr = (int)cpr;
  f->R = R;
  f->cells = (int)(cpr * R);
  f->nsb = (f->cells + 256 * kScanItems - 1) / (256 * kScanItems);
  f->st = st;
  // replica offsets and the coarse position -> replica table
  f->rep_first_h.assign(R + 1, 0);
  for (int r = 0; r < R; ++r) f->rep_first_h[r + 1] = f->rep_first_h[r] + rep_n[r];
  if (f->rep_first_h[R] != P.n) {
    delete f;
    g_fast_err = "replica sizes do not add up to n";
    return SPSS_ERR_INVALID;
  }
  const int nchunk = (P.n >> kChunkShift) + 1;
  std::vector<int> chunk(nchunk, 0);
  for (int q = 0, r = 0; q < nchunk; ++q) {
    const int p = q << kChunkShift;
    while (r + 1 < R && f->rep_first_h[r + 1] <= p) ++r;
    chunk[q] = r;
  }
  const int cells = f->cells;
  const size_t n = (size_t)P.n;
  hipError_t e = hipSuccess;
  auto A = [&](hipError_t r) {
    if (e == hipSuccess) e = r;
  };
  A(falloc(&f->key, n));
  A(falloc(&f->slot, n));
  A(falloc(&f->count, cells));
  A(falloc(&f->start, cells + 1));
  A(falloc(&f->tmp, n));
  A(falloc(&f->sorted, n));
  A(falloc(&f->blocksum, f->nsb + 1));
  if (R == 1 && f->ncx <= kRowTileCols) {
    // slots per row: the greedy spans of a row 1.5x the mean, at about 94% of kFB each, plus 2,
    // and at least enough spans of rt_wcap columns to cover a row
    const double row_mean = (double)P.n / (double)f->ncy;
    f->rt_wcap = std::min(64, max_tile_width(f));
    f->rt_cmax = std::max((f->ncx + f->rt_wcap - 1) / f->rt_wcap,
                          std::min(f->ncx, (int)std::ceil(1.5 * row_mean / (0.94 * kFB)) + 2));
    A(falloc(&f->rtiles, (size_t)f->ncy * f->rt_cmax));
    A(falloc(&f->btiles, (size_t)f->ncy * f->rt_cmax));
  }
  A(falloc(&f->snap, n));
  A(falloc(&f->omega_f, n));
  for (int b = 0; b < 2; ++b) {
    A(falloc(&f->rstate[b], n * P.s));
    A(falloc(&f->rid[b], n));
    A(falloc(&f->rnat[b], n));
    A(falloc(&f->rstart[b], cells + 1));
  }
  A(falloc(&f->rsnap, n));
  A(falloc(&f->romega, n));
  A(falloc(&f->ssum, n));
  A(falloc(&f->tsum, n));
  A(falloc(&f->scnt, n));
  A(falloc(&f->mvb[0], n));
  A(falloc(&f->mvb[1], n));
  A(falloc(&f->cnt9[0], (size_t)cells * 9));
  A(falloc(&f->cnt9[1], (size_t)cells * 9));
  A(falloc(&f->farcnt, cells));
  A(falloc(&f->newcnt, cells));
  A(falloc(&f->rep_first, R + 1));
  A(falloc(&f->ovf_count, 1));
  // overflow work: at most one span per cell, or one item per cell and kFB particles
  A(falloc(&f->ovf_list, (size_t)cells + n / kFB + 1));
  A(falloc(&f->rep_chunk, nchunk));
  A(falloc(&f->step_dev, 1));
  A(falloc(&f->lb_status, f->nsb));
  for (int b = 0; b < 2; ++b) {
    A(falloc(&f->farlist[b], 2L * n));  // {slot, destination cell} per far mover
    A(falloc(&f->farlist_n[b], 1));
  }
  A(falloc(&f->farslot, n));
  A(falloc(&f->lb_error, 1));
  if (coef) {
    A(falloc(&f->coef, R));
    A(falloc(&f->zlast, 2 * R));
    A(falloc(&f->racc, 3 * R));
  }
  if (e == hipSuccess) e = hipMemcpy(f->rep_first, f->rep_first_h.data(), sizeof(int) * (R + 1), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(f->rep_chunk, chunk.data(), sizeof(int) * nchunk, hipMemcpyHostToDevice);
  if (e == hipSuccess && coef) e = hipMemcpy(f->coef, coef, sizeof(Coef) * R, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    fast_destroy(f);
    return ffail(std::string("hipMalloc/hipMemcpy: ") + hipGetErrorString(e));
  }
  std::vector<int> W(R);
  for (int r = 0; r < R; ++r) W[r] = default_tile_width(f, (double)rep_n[r] / (double)cpr);
  int rc = set_tiles(f, W);
  if (rc != SPSS_OK) {
    fast_destroy(f);
    return rc;
  }
  launch_fill_int(st, cells, f->count, 0);
  launch_fill_int(st, cells, f->farcnt, 0);
  launch_fill_int(st, (long)cells * 9, f->cnt9[0], 0);
  launch_fill_int(st, (long)cells * 9, f->cnt9[1], 0);
  launch_fill_int(st, cells, f->newcnt, 0);
  launch_fill_int(st, 1, f->ovf_count, 0);
  launch_fill_int(st, 2, reinterpret_cast<int *>(f->step_dev), 0);  // the 64-bit step counter
  launch_fill_int(st, 2L * f->nsb, reinterpret_cast<int *>(f->lb_status), 0);
  launch_fill_int(st, 1, f->farlist_n[0], 0);
  launch_fill_int(st, 1, f->farlist_n[1], 0);
  launch_fill_int(st, 1, f->lb_error, 0);
  if (coef) {
    launch_fill_double(st, 3L * R, f->racc, 0.0);
    launch_fill_double(st, 2L * R, f->zlast, 0.0);
  }
  *out = f;
  return SPSS_OK;
}

int fast_create(FastPath **out, hipStream_t st, const DevParams &P, double rho, int nc, int subdiv, int xsplit) {
  const int n = P.n;
  return fast_create_replicas(out, st, P, rho, nc, subdiv, xsplit, 1, &n, nullptr);
}

int fast_cell_xsplit(const FastPath *f) { return f ? f->xsplit : 1; }

void fast_destroy(FastPath *f) {
  if (!f) return;
  for (auto &row : f->sg)
    for (auto &g : row) {
      if (g.exec) hipGraphExecDestroy(g.exec);
      if (g.graph) hipGraphDestroy(g.graph);
    }
  void *b[] = {f->key,       f->slot,      f->count,     f->start,   f->tmp,       f->sorted,    f->blocksum,
               f->snap,      f->omega_f,   f->rstate[0], f->rstate[1], f->rid[0],  f->rid[1],    f->rnat[0],
               f->rnat[1],   f->rstart[0], f->rstart[1], f->rsnap,   f->romega,    f->newcnt,    f->mvb[0], f->mvb[1],
               f->cnt9[0],   f->cnt9[1],   f->farcnt,    f->ssum,      f->tsum,    f->scnt,      f->rep_first, f->rep_chunk,
               f->tiles,     f->coef,      f->step_dev,  f->zlast,   f->racc,      f->ovf_list,  f->ovf_count,
               f->rtiles,    f->btiles,    f->lb_status, f->farlist[0], f->farlist[1], f->farlist_n[0],
               f->farlist_n[1], f->lb_error, f->farslot};
  for (void *p : b) hipFree(p);
  delete f;
}

int fast_tile_widths(FastPath *f, std::vector<int> *W) {
  *W = f->W_h;
  return SPSS_OK;
}

int fast_set_tile_widths(FastPath *f, const std::vector<int> &W) {
  if ((int)W.size() != f->R) {
    g_fast_err = "one tile width per replica";
    return SPSS_ERR_INVALID;
  }
  return set_tiles(f, W);
}

void fast_set_stream(FastPath *f, hipStream_t st) {
  if (f) f->st = st;
}

void fast_set_timer(FastPath *f, KernelTimer *t) {
  if (f) f->timer = t;
}

// the XCD-aware tile order needs a grid that is a multiple of the 8 XCDs
static int tile_grid(int ntiles) { return ((ntiles + 7) / 8) * 8; }

static unsigned nblocks(long n) { return (unsigned)((n + 255) / 256); }

// workgroups of the overflow pass (grid-stride over its items): the dense rows of a clustered
// replica batch give thousands of items (512 clustered replicas: 5.9e9 particle-steps/s with 4096
// workgroups, 5.7e9 with 1024, 3.8e9 with 256); a single system's tiles rarely overflow, and the
// pass then costs its dispatch, which grows with the grid
constexpr int kOverflowBlocks = 4096;
constexpr int kOverflowBlocksSingle = 512;
static int overflow_blocks(const FastPath *f) { return f->R > 1 ? kOverflowBlocks : kOverflowBlocksSingle; }

static FastArgs args_of(const FastPath *f) {
  FastArgs a;
  a.n = f->n;
  a.ncx = f->ncx;
  a.ncy = f->ncy;
  a.kx = f->kx;
  a.ky = f->ky;
  a.ntiles = f->ntiles;
  a.cpr = f->cpr;
  a.tiles = f->tiles;
  a.ovf_list = f->ovf_list;
  a.ovf_count = f->ovf_count;
  a.Hx = f->Hx;
  a.Hy = f->Hy;
  a.r2c = f->r2c;
  a.r2h = f->r2h;
  a.r2k = f->r2k;
  return a;
}

static Layout layout_of(const FastPath *f) {
  Layout L;
  L.rep_first = f->rep_first;
  L.rep_chunk = f->rep_chunk;
  L.cpr = f->cpr;
  L.nrep = f->R;
  return L;
}

// sorted order, cell starts and snapshot of a stage state (original order)
static void build(FastPath *f, hipStream_t st, const DevParams &P, const double *stage, int *start, float4 *snap,
                  float *omega) {
  const int n = f->n;
  const unsigned nb = nblocks(n);
  const Layout L = layout_of(f);
  if (P.s == 4)
    fk_keys<4><<<nb, 256, 0, st>>>(P, L, stage, f->ncx, f->ncy, f->key, f->slot, f->count);
  else
    fk_keys<3><<<nb, 256, 0, st>>>(P, L, stage, f->ncx, f->ncy, f->key, f->slot, f->count);
  scan_local<<<f->nsb, 256, 0, st>>>(f->count, nullptr, f->cells, start, f->blocksum);
  scan_add<<<f->nsb, 256, 0, st>>>(f->count, nullptr, f->cells, start, f->blocksum, f->nsb, f->ovf_count);
  fk_scatter<<<nb, 256, 0, st>>>(n, f->key, f->slot, start, f->tmp);
  if (P.s == 4)
    fk_rank<4><<<nb, 256, 0, st>>>(P, stage, f->ncx, f->cpr, f->Hx, f->Hy, f->key, start, f->tmp, f->sorted, snap,
                                   omega);
  else
    fk_rank<3><<<nb, 256, 0, st>>>(P, stage, f->ncx, f->cpr, f->Hx, f->Hy, f->key, start, f->tmp, f->sorted, snap,
                                   nullptr);
}

// row tiles of the rebuilt cell list (single systems), from the starts build() left in f->start
static void rebuilt_row_tiles(FastPath *f, hipStream_t st) {
  if (f->btiles)
    fk_row_tiles<<<(f->ncy + 3) / 4, 256, 0, st>>>(f->start, f->ncx, f->ncy, f->rt_wcap, f->rt_cmax, f->btiles);
}

static PairBufs rebuilt_bufs(const FastPath *f, const double *stage) {
  PairBufs B{};
  B.stage = stage;
  B.sorted = f->sorted;
  B.start = f->start;
  B.snap = f->snap;
  B.omega_f = f->omega_f;
  return B;
}

// the pair kernel of a rebuilt cell list (MODE 0-2), then the pass over the tiles whose cells
// overflowed its LDS stage (usually none); the resident step has resident_pairs
template <int S, bool SRK, int MODE>
static void launch_pairs(const FastPath *f, hipStream_t st, const DevParams &P, const PairBufs &B) {
  static_assert(MODE != 3, "resident steps launch fr_pairs");
  FastArgs a = args_of(f);
  if (f->btiles) {  // the rebuilt cell list's row tiles (rebuilt_row_tiles)
    a.tiles = f->btiles;
    a.ntiles = f->ncy * f->rt_cmax;
  }
  fk_pairs<S, SRK, MODE><<<tile_grid(a.ntiles), kFB, 0, st>>>(P, a, B);
  fk_pairs_ovf<S, SRK, MODE><<<overflow_blocks(f), kFB, 0, st>>>(P, a, B);
}

int fast_rhs(FastPath *f, hipStream_t st, const DevParams &P, const double *stage, const double *nat,
             double *k_next, double *compl_out) {
  const bool srk = P.higher_order && compl_out && P.s == 4;
  build(f, st, P, stage, f->start, f->snap, srk ? f->omega_f : nullptr);
  rebuilt_row_tiles(f, st);
  PairBufs B = rebuilt_bufs(f, stage);
  B.nat = nat;
  B.k_next = k_next;
  B.compl_out = srk ? compl_out : nullptr;
  if (f->timer) f->timer->begin(st);
  if (P.s == 4 && srk)
    launch_pairs<4, true, 0>(f, st, P, B);
  else if (P.s == 4)
    launch_pairs<4, false, 0>(f, st, P, B);
  else
    launch_pairs<3, false, 0>(f, st, P, B);
  if (f->timer) f->timer->end(st);
  FCHK(hipGetLastError());
  return SPSS_OK;
}

bool fast_is_resident(const FastPath *f) { return f && f->resident; }

int fast_check(FastPath *f, hipStream_t st) {
  int err = 0;
  FCHK(hipMemcpyAsync(&err, f->lb_error, sizeof(int), hipMemcpyDeviceToHost, st));
  FCHK(hipStreamSynchronize(st));
  if (err) return ffail("a look-back wait of the arrivals scan ran out: the cell starts of some step are not valid");
  return SPSS_OK;
}

int fast_buffer_parity(const FastPath *f) { return f ? 2 * f->cur + f->mvc : -1; }

void fast_resident_drop(FastPath *f) {
  if (f) f->resident = false;
}

int fast_resident_sync(FastPath *f, hipStream_t st, double *state) {
  if (!f->resident) return SPSS_OK;
  const unsigned nb = nblocks(f->n);
  if (f->s == 4)
    fr_unpermute<4><<<nb, 256, 0, st>>>(f->n, f->rstate[f->cur], f->rid[f->cur], state);
  else
    fr_unpermute<3><<<nb, 256, 0, st>>>(f->n, f->rstate[f->cur], f->rid[f->cur], state);
  FCHK(hipGetLastError());
  return SPSS_OK;
}

// forget the moves the last fr_scatter computed: their counts go back to zero
static void drop_moves(FastPath *f, hipStream_t st) {
  if (!f->mv_ready) return;
  launch_fill_int(st, f->cells, f->newcnt, 0);
  launch_fill_int(st, f->cells, f->farcnt, 0);
  launch_fill_int(st, (long)f->cells * 9, f->cnt9[f->cur], 0);
  launch_fill_int(st, 1, f->farlist_n[f->cur], 0);
  f->mv_ready = false;
}

// the moves and counts of the current step: from the last fr_scatter when it computed them
// for this dt, stepper and v0, else from fr_move
template <int S, bool SRK>
static void ensure_moves(FastPath *f, hipStream_t st, const DevParams &P, const StepArgs &U) {
  if (f->mv_ready && f->mv_dt == U.dt && f->mv_srk == SRK && f->mv_v0 == P.v0) return;
  drop_moves(f, st);
  fr_move<S, SRK><<<nblocks(f->n), 256, 0, st>>>(P, layout_of(f), f->n, f->ncx, f->ncy, f->rstate[f->cur], U,
                                                 f->mvb[f->mvc], f->farcnt, f->farslot, f->cnt9[f->cur], f->newcnt,
                                                 f->farlist[f->cur], f->farlist_n[f->cur]);
  f->mv_ready = true;
  f->mv_dt = U.dt;
  f->mv_srk = SRK;
  f->mv_v0 = P.v0;
}

static void ensure_moves_of(FastPath *f, hipStream_t st, const DevParams &P, bool do_srk, const StepArgs &U) {
  if (do_srk)
    ensure_moves<4, true>(f, st, P, U);
  else if (P.s == 4)
    ensure_moves<4, false>(f, st, P, U);
  else
    ensure_moves<3, false>(f, st, P, U);
}

// entry into resident mode: cell order of `state`, then the gather into the resident arrays
static void resident_enter(FastPath *f, hipStream_t st, const DevParams &P, const double *state,
                           const double *nat, bool srk) {
  const int c = f->cur;
  build(f, st, P, state, f->rstart[c], f->rsnap, (srk && P.s == 4) ? f->romega : nullptr);
  const unsigned nb = nblocks(f->n);
  if (P.s == 4)
    fr_gather<4><<<nb, 256, 0, st>>>(f->n, state, f->sorted, nat, f->rstate[c], f->rid[c], f->rnat[c]);
  else
    fr_gather<3><<<nb, 256, 0, st>>>(f->n, state, f->sorted, nat, f->rstate[c], f->rid[c], f->rnat[c]);
  if (f->rtiles)
    fk_row_tiles<<<(f->ncy + 3) / 4, 256, 0, st>>>(f->rstart[c], f->ncx, f->ncy, f->rt_wcap, f->rt_cmax, f->rtiles);
  drop_moves(f, st);
  f->resident = true;
  f->resident_nat = nat != nullptr;
}

// the arrivals of the current step (its moves are in place: ensure_moves)
static Arrivals arrivals_of(const FastPath *f) {
  const int c = f->cur, o = 1 - c;
  Arrivals Q;
  Q.newcnt = f->newcnt;
  Q.farcnt = f->farcnt;
  Q.m = f->cells;
  Q.nsb = f->nsb;
  Q.nstart = f->rstart[o];
  Q.status = f->lb_status;
  Q.farlist = f->farlist[c];
  Q.farlist_n = f->farlist_n[c];
  Q.farlist_n_next = f->farlist_n[o];
  Q.cnt9 = f->cnt9[c];
  Q.farslot = f->farslot;
  Q.bucket = f->tmp;
  Q.cnt9_next = f->cnt9[o];
  Q.nclear = (long)f->cells * 9;
  Q.ncx = f->ncx;
  Q.ncy = f->ncy;
  Q.cpr = f->cpr;
  Q.n = f->n;
  Q.error = f->lb_error;
  return Q;
}

// the pair kernel of a resident step with the step's arrival scan in its first workgroups
// (fr_pairs), then the far arrivals and the work of tiles that overflowed the LDS stage (fr_far,
// timed apart)
template <int S, bool SRK>
static void resident_pairs(const FastPath *f, hipStream_t st, const DevParams &P, const StepArgs &U,
                           const PairBufs &B) {
  FastArgs a = args_of(f);
  if (f->rtiles) {  // the resident order's row tiles, rebuilt every step
    a.tiles = f->rtiles;
    a.ntiles = f->ncy * f->rt_cmax;
  }
  const int nsb8 = (f->nsb + 7) & ~7;
  const Arrivals Q = arrivals_of(f);
  if (f->timer) f->timer->begin(st);
  fr_pairs<S, SRK><<<nsb8 + tile_grid(a.ntiles), kFB, 0, st>>>(P, a, B, Q, U);
  if (f->timer) f->timer->end(st);
  fr_far<S, SRK><<<overflow_blocks(f), kFB, 0, st>>>(P, a, B, Q);
}

// the step applied and written to the new slots (after resident_pairs)
template <int S, bool SRK>
static void resident_scatter(FastPath *f, hipStream_t st, const DevParams &P, const StepArgs &U, double *compl_out) {
  const int c = f->cur, o = 1 - c;
  const int n = f->n;
  const unsigned nb = nblocks(n);
  const double *rnat = f->resident_nat ? f->rnat[c] : nullptr;
  const Layout L = layout_of(f);
  int *const mv = f->mvb[f->mvc], *const mvn = f->mvb[1 - f->mvc];
  // (at least one workgroup per four rows of cells for the row tiles)
  const unsigned nbs = std::max(nb, (unsigned)(f->rtiles ? (f->ncy + 3) / 4 : 0));
  fr_scatter<S, SRK><<<nbs, 256, 0, st>>>(P, L, n, f->ncx, f->ncy, f->Hx, f->Hy, f->rstart[c], f->rstart[o], mv,
                                        f->cnt9[c], f->tmp,
                                        f->rstate[c], f->rid[c], rnat, f->ssum, f->tsum, f->scnt, U,
                                        SRK ? compl_out : nullptr, f->rstate[o], f->rid[o], f->rnat[o], f->rsnap,
                                        SRK ? f->romega : nullptr, f->rtiles, f->rt_wcap, f->rt_cmax, mvn, f->farcnt,
                                        f->farslot, f->cnt9[o], f->newcnt, f->farlist[o], f->farlist_n[o],
                                        f->ovf_count);
  f->cur = o;
  f->mvc = 1 - f->mvc;  // the next step's moves (same dt, stepper and v0) are in place
}

static int step_impl(FastPath *f, hipStream_t st, const DevParams &P, bool do_srk, const double *state,
                     const double *nat, double *compl_out, const StepArgs &U) {
  // the SRK step needs the FP32 omega of every particle, which an Euler-entered copy lacks
  if (f->resident && (nat != nullptr) != f->resident_nat) f->resident = false;
  if (f->resident && do_srk && !f->romega_valid) f->resident = false;
  if (!f->resident) resident_enter(f, st, P, state, nat, do_srk);
  f->romega_valid = do_srk;
  const int c = f->cur;
  PairBufs B{};
  B.rstate = f->rstate[c];
  B.rid = f->rid[c];
  B.start = f->rstart[c];
  B.snap = f->rsnap;
  B.omega_f = f->romega;
  B.ssum = f->ssum;
  B.tsum = f->tsum;
  B.scnt = f->scnt;
  ensure_moves_of(f, st, P, do_srk, U);  // the arrivals in the pair launch scan them
  if (do_srk) {
    resident_pairs<4, true>(f, st, P, U, B);
    resident_scatter<4, true>(f, st, P, U, compl_out);
  } else if (P.s == 4) {
    resident_pairs<4, false>(f, st, P, U, B);
    resident_scatter<4, false>(f, st, P, U, nullptr);
  } else {
    resident_pairs<3, false>(f, st, P, U, B);
    resident_scatter<3, false>(f, st, P, U, nullptr);
  }
  FCHK(hipGetLastError());
  return SPSS_OK;
}

static StepArgs step_args(double dt, double D_phi, double friction, int noise_mode, const double *noise,
                          uint64_t seed, uint64_t step) {
  StepArgs U{};
  U.dt = dt;
  U.sqrt_dt = std::sqrt(dt);
  U.int_phi = std::sqrt(2.0 * D_phi);
  U.c15 = 0.5 * std::pow(dt, 1.5);
  U.is3 = 1.0 / std::sqrt(3.0);
  U.sq2d = std::sqrt(2.0 * D_phi);
  U.friction = friction;
  U.noise_mode = noise_mode;
  U.noise = noise;
  U.seed = seed;
  U.step = step;
  return U;
}

int fast_step(FastPath *f, hipStream_t st, const DevParams &P, int srk, const double *state, const double *nat,
              double *compl_out, double dt, double D_phi, double friction, int noise_mode, const double *noise,
              uint64_t seed, uint64_t step) {
  const StepArgs U = step_args(dt, D_phi, friction, noise_mode, noise, seed, step);
  return step_impl(f, st, P, srk && P.s == 4, state, nat, compl_out, U);
}

// ---------------------------------------------------------------------------
// graph-replayed fused steps of a single system.  Each captured step advances the device step
// counter in its fr_pairs and reads it as its Philox step in its fr_scatter, so one graph
// serves every replay.  Graphs of kGraphSteps steps (even: the double buffers end where they
// started) carry the bulk of a run, graphs of one step the remainder; a graph is replayed only
// from the buffer parity it was captured at.  Capture records launches without running them, so
// the host bookkeeping (buffer parity, move flags) is put back after every capture and advanced
// by each replay instead.
static constexpr int kGraphSteps = 10;

__global__ void fs_set_counter(unsigned long long *c, unsigned long long v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *c = v;
}

static void drop_step_graphs(FastPath *f) {
  for (auto &row : f->sg)
    for (auto &g : row) {
      if (g.exec) hipGraphExecDestroy(g.exec);
      if (g.graph) hipGraphDestroy(g.graph);
      g.exec = nullptr;
      g.graph = nullptr;
    }
}

FastHostMark fast_host_mark(const FastPath *f) {
  return FastHostMark{f->resident, f->resident_nat, f->romega_valid, f->mv_ready, f->mv_srk,
                      f->cur,      f->mvc,          f->mv_dt,        f->mv_v0};
}

void fast_restore_host_mark(FastPath *f, const FastHostMark &m) {
  f->resident = m.resident;
  f->resident_nat = m.resident_nat;
  f->romega_valid = m.romega_valid;
  f->mv_ready = m.mv_ready;
  f->mv_srk = m.mv_srk;
  f->cur = m.cur;
  f->mvc = m.mvc;
  f->mv_dt = m.mv_dt;
  f->mv_v0 = m.mv_v0;
}

// one replayed (or skipped-over) step moves both double buffers
static void flip_parity(FastPath *f) {
  f->cur = 1 - f->cur;
  f->mvc = 1 - f->mvc;
}

bool fast_graph_ready(const FastPath *f, const double *nat, int srk) {
  return f && !f->timer && f->resident && (nat != nullptr) == f->resident_nat && (!srk || f->romega_valid);
}

// records `steps` fused steps from the current parity into g (host bookkeeping restored after)
static int capture_steps(FastPath *f, hipStream_t st, const DevParams &P, bool do_srk, const double *state,
                         const double *nat, double *compl_out, const StepArgs &U, int steps,
                         FastPath::StepGraph &g) {
  const FastHostMark m = fast_host_mark(f);
  hipError_t e = hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed);
  if (e != hipSuccess) return ffail(std::string("hipStreamBeginCapture: ") + hipGetErrorString(e));
  int rc = SPSS_OK;
  for (int k = 0; k < steps && rc == SPSS_OK; ++k) rc = step_impl(f, st, P, do_srk, state, nat, compl_out, U);
  hipGraph_t graph = nullptr;
  e = hipStreamEndCapture(st, &graph);
  fast_restore_host_mark(f, m);  // nothing ran
  if (rc != SPSS_OK || e != hipSuccess) {
    if (graph) hipGraphDestroy(graph);
    (void)hipGetLastError();
    if (rc == SPSS_OK) return ffail(std::string("capture of the step graph failed: ") + hipGetErrorString(e));
    return rc;
  }
  hipGraphExec_t exec = nullptr;
  e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    hipGraphDestroy(graph);
    return ffail(std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
  }
  g.graph = graph;
  g.exec = exec;
  ++f->sg_captures;
  return SPSS_OK;
}

// the moves of the first step in place (never from a captured fr_move: a replay after a step that
// left its moves would count them twice), and the graphs dropped when the key changed
static StepArgs graph_setup(FastPath *f, hipStream_t st, const DevParams &P, bool do_srk, const double *state,
                            const double *nat, double *compl_out, double dt, double D_phi, double friction,
                            uint64_t seed) {
  ensure_moves_of(f, st, P, do_srk, step_args(dt, D_phi, friction, 3, nullptr, seed, 0));
  const bool same = f->sg_srk == (int)do_srk && f->sg_gen == f->gen && f->sg_higher == P.higher_order &&
                    f->sg_dt == dt && f->sg_dphi == D_phi && f->sg_friction == friction && f->sg_seed == seed &&
                    f->sg_state == state && f->sg_nat == nat && f->sg_compl == compl_out;
  if (!same) {
    drop_step_graphs(f);
    f->sg_srk = do_srk;
    f->sg_gen = f->gen;
    f->sg_higher = P.higher_order;
    f->sg_dt = dt;
    f->sg_dphi = D_phi;
    f->sg_friction = friction;
    f->sg_seed = seed;
    f->sg_state = state;
    f->sg_nat = nat;
    f->sg_compl = compl_out;
  }
  StepArgs U = step_args(dt, D_phi, friction, 3, nullptr, seed, 0);
  U.step_dev = f->step_dev;
  U.step_inc = f->step_dev;
  return U;
}

static int ensure_graph(FastPath *f, hipStream_t st, const DevParams &P, bool do_srk, const double *state,
                        const double *nat, double *compl_out, const StepArgs &U, int size) {
  FastPath::StepGraph &g = f->sg[size == 1 ? 1 : 0][fast_buffer_parity(f)];
  if (g.exec) return SPSS_OK;
  return capture_steps(f, st, P, do_srk, state, nat, compl_out, U, size, g);
}

int fast_enter(FastPath *f, hipStream_t st, const DevParams &P, int srk, const double *state, const double *nat) {
  const bool do_srk = srk && P.s == 4;
  if (fast_graph_ready(f, nat, do_srk)) return SPSS_OK;
  if (f->timer) {
    g_fast_err = "the step graphs are not captured while the kernel timer is on";
    return SPSS_ERR_INVALID;
  }
  resident_enter(f, st, P, state, nat, do_srk);
  f->romega_valid = do_srk;
  FCHK(hipGetLastError());
  return SPSS_OK;
}

int fast_prepare_graphs(FastPath *f, hipStream_t st, const DevParams &P, int srk, const double *state,
                        const double *nat, double *compl_out, double dt, double D_phi, double friction,
                        uint64_t seed) {
  const bool do_srk = srk && P.s == 4;
  if (!fast_graph_ready(f, nat, do_srk)) {
    g_fast_err = "the step graphs need the resident order in this stepper's form (fast_enter)";
    return SPSS_ERR_INVALID;
  }
  const StepArgs U = graph_setup(f, st, P, do_srk, state, nat, compl_out, dt, D_phi, friction, seed);
  // from this parity and from the other one (the host is set there for the capture only)
  for (int side = 0; side < 2; ++side) {
    for (int size : {kGraphSteps, 1}) {
      const int rc = ensure_graph(f, st, P, do_srk, state, nat, compl_out, U, size);
      if (rc != SPSS_OK) {
        if (side) flip_parity(f);
        return rc;
      }
    }
    flip_parity(f);
  }
  return SPSS_OK;
}

long fast_run_graph(FastPath *f, hipStream_t st, const DevParams &P, int srk, const double *state, const double *nat,
                    double *compl_out, double dt, double D_phi, double friction, uint64_t seed, uint64_t step,
                    long nsteps) {
  const bool do_srk = srk && P.s == 4;
  // only from the resident order in the form this stepper needs, and not while the kernel timer
  // records events around the pair kernel
  if (nsteps < 1 || !fast_graph_ready(f, nat, do_srk)) return 0;
  const StepArgs U = graph_setup(f, st, P, do_srk, state, nat, compl_out, dt, D_phi, friction, seed);
  const long reps = nsteps / kGraphSteps, rest = nsteps % kGraphSteps;
  // every graph this run replays, captured before anything is launched
  int rc = SPSS_OK;
  if (reps > 0) rc = ensure_graph(f, st, P, do_srk, state, nat, compl_out, U, kGraphSteps);
  for (int side = 0; side < 2 && side < rest && rc == SPSS_OK; ++side) {
    if (side) flip_parity(f);
    rc = ensure_graph(f, st, P, do_srk, state, nat, compl_out, U, 1);
    if (side) flip_parity(f);
  }
  if (rc != SPSS_OK) return -1;
  auto failed = [&](const char *what, hipError_t e) -> long {
    g_fast_err = std::string(what) + ": " + hipGetErrorString(e);
    return -1;
  };
  fs_set_counter<<<1, 64, 0, st>>>(f->step_dev, (unsigned long long)step - 1ull);  // advanced before use
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return failed("fs_set_counter", e);
  for (long r = 0; r < reps; ++r)
    if ((e = hipGraphLaunch(f->sg[0][fast_buffer_parity(f)].exec, st)) != hipSuccess) return failed("hipGraphLaunch", e);
  for (long r = 0; r < rest; ++r) {
    if ((e = hipGraphLaunch(f->sg[1][fast_buffer_parity(f)].exec, st)) != hipSuccess) return failed("hipGraphLaunch", e);
    flip_parity(f);
  }
  f->romega_valid = do_srk;
  return nsteps;
}

long fast_graph_captures(const FastPath *f) { return f ? f->sg_captures : 0; }

// ---------------------------------------------------------------------------
// replica batches: per-replica coefficients, Philox step from the device counter

int fast_batch_step(FastPath *f, hipStream_t st, const DevParams &P, const double *state, const double *nat,
                    double dt) {
  if (!f->coef) {
    g_fast_err = "not a replica batch";
    return SPSS_ERR_INVALID;
  }
  StepArgs U = step_args(dt, 0.0, 0.0, 3, nullptr, 0, 0);
  U.coef = f->coef;
  U.step_dev = f->step_dev;
  int rc = step_impl(f, st, P, false, state, nat, nullptr, U);
  if (rc != SPSS_OK) return rc;
  fb_step_inc<<<1, 64, 0, st>>>(f->step_dev);
  FCHK(hipGetLastError());
  return SPSS_OK;
}

int fast_prepare_moves(FastPath *f, hipStream_t st, const DevParams &P, double dt) {
  if (!f->resident) return SPSS_OK;
  ensure_moves_of(f, st, P, false, step_args(dt, 0.0, 0.0, 3, nullptr, 0, 0));
  FCHK(hipGetLastError());
  return SPSS_OK;
}

int fast_batch_observe(FastPath *f, hipStream_t st, uint64_t from) {
  fb_order_stats<<<f->R, 256, 0, st>>>(layout_of(f), f->rsnap, f->step_dev, (unsigned long long)from, f->zlast,
                                       f->racc);
  FCHK(hipGetLastError());
  return SPSS_OK;
}

int fast_batch_reset(FastPath *f, hipStream_t st, uint64_t step) {
  launch_fill_double(st, 3L * f->R, f->racc, 0.0);
  const unsigned long long v = step;
  FCHK(hipMemcpyAsync(f->step_dev, &v, sizeof(v), hipMemcpyHostToDevice, st));
  FCHK(hipStreamSynchronize(st));
  return SPSS_OK;
}

int fast_batch_stats(FastPath *f, hipStream_t st, double *racc, double *zlast, uint64_t *step) {
  unsigned long long v = 0;
  if (racc) FCHK(hipMemcpyAsync(racc, f->racc, sizeof(double) * 3 * f->R, hipMemcpyDeviceToHost, st));
  if (zlast) FCHK(hipMemcpyAsync(zlast, f->zlast, sizeof(double) * 2 * f->R, hipMemcpyDeviceToHost, st));
  FCHK(hipMemcpyAsync(&v, f->step_dev, sizeof(v), hipMemcpyDeviceToHost, st));
  FCHK(hipStreamSynchronize(st));
  if (step) *step = v;
  return SPSS_OK;
}

int fast_replica_count(const FastPath *f) { return f->R; }

int fast_pairs_count(FastPath *f, hipStream_t st, const DevParams &P, const double *stage, int *count) {
  build(f, st, P, stage, f->start, f->snap, nullptr);
  rebuilt_row_tiles(f, st);
  PairBufs B = rebuilt_bufs(f, stage);
  B.pcount = count;
  if (P.s == 4)
    launch_pairs<4, false, 1>(f, st, P, B);
  else
    launch_pairs<3, false, 1>(f, st, P, B);
  FCHK(hipGetLastError());
  return SPSS_OK;
}

// lists of the stage state of the preceding fast_pairs_count
int fast_pairs_fill(FastPath *f, hipStream_t st, const DevParams &P, const double *stage, const int *pstart,
                    int *plist) {
  launch_fill_int(st, 1, f->ovf_count, 0);  // the count pass left its overflow list behind
  PairBufs B = rebuilt_bufs(f, stage);
  B.pstart = pstart;
  B.plist = plist;
  if (P.s == 4)
    launch_pairs<4, false, 2>(f, st, P, B);
  else
    launch_pairs<3, false, 2>(f, st, P, B);
  FCHK(hipGetLastError());
  return SPSS_OK;
}

}  // namespace spss
