// hydro_fused.h -- the face-fused gfx950 engine of the hydro advance.
#pragma once

#include <vector>

#include "canon.h"
#include "hydro_state.h"

namespace camr {


// hydro_ctoprim (Source/Hydro/Hydro_ctoprim.H:11-101) of the conserved
// state v (NVAR values) into the NQC compact primitives o; e receives the
// specific internal energy (needed by hydro_srctoprim).  Returns false and
// leaves o untouched for rho <= 0, as the reference leaves such cells.
__device__ __forceinline__ bool prim_compact(const double *v, double *o, double &e,
                                             const HydroConst &hc, int *status) {
    const double rho = v[URHO];
    if (!(rho > 0.)) return false;
    const double rhoinv = 1.0 / rho;
    const double vx = v[UMX] * rhoinv;
    const double vy = v[UMY] * rhoinv;
    const double vz = v[UMZ] * rhoinv;
    const double kineng = 0.5 * rho * (vx * vx + vy * vy + vz * vz);
    const double ueden = v[UEDEN];
    const double ueint = v[UEINT];
    if (!hc.allow_negative_energy && !(ueden > 0. && ueint > 0.))
        atomicOr(status, 1 << CAMR_ERR_NEG_ENERGY);
    e = ((ueden - kineng) / ueden > hc.eta1) ? (ueden - kineng) * rhoinv : ueint * rhoinv;
    const double p = (kGamma - 1.0) * rho * e;
    const double cs = sqrt(kGamma * p / rho);
    o[SR] = rho;
    o[SU] = vx;
    o[SV] = vy;
    o[SW] = vz;
    o[SP] = p;
    o[SE] = e * rho;
#pragma unroll
    for (int s = 0; s < NSPEC; ++s) o[SY + s] = v[UFS + s] * rhoinv;
    o[SC] = cs;
    o[SCS] = vmax(hc.small_num, hc.small_num * cs);
    return true;
}

// The same, stored at (i,j,k) of q.
__device__ __forceinline__ void prim_compact_point(const double *v, const Arr &q, int i,
                                                   int j, int k, const HydroConst &hc,
                                                   int *status) {
    double o[NQC], e;
    if (!prim_compact(v, o, e, hc, status)) return;
#pragma unroll
    for (int n = 0; n < NQC; ++n) q(i, j, k, n) = o[n];
}

// ------------------------------------------------------------ box batches
// The engine runs every stage either over all boxes of the level in one
// launch, or over one box per launch; both use the same kernels.  A block
// finds its box by a binary search of the stage's block prefix table.

// Index ranges of the stages (one prefix table per kind).
enum StageKind {
    ST_G4 = 0,        // grow(bx, 4): ghost fill, primitives
    ST_G2 = 1,        // grow(bx, 2): flattening, traces
    ST_R1 = 2,        // +D: faces of the first Riemann solves
    ST_T1 = 5,        // +D: faces of the transdo pairs
    ST_DIVU = 8,      // nodes of bx
    ST_T2 = 9,        // +D: faces of bx
    ST_VALID = 12,    // bx
    ST_COUNT = 13
};

__host__ __device__ inline Box stage_box(int kind, const Box &bx) {
    if (kind == ST_G4) return grow(bx, CAMR_NGROW);
    if (kind == ST_G2) return grow(bx, 2);
    if (kind >= ST_R1 && kind < ST_R1 + 3) {
        const int d = kind - ST_R1;
        return nodes(grow_dir(grow(bx, 2), d, -1), d);
    }
    if (kind >= ST_T1 && kind < ST_T1 + 3) return nodes(grow(bx, 1), kind - ST_T1);
    if (kind == ST_DIVU) return nodes(nodes(nodes(bx, 0), 1), 2);
    if (kind >= ST_T2 && kind < ST_T2 + 3) return nodes(bx, kind - ST_T2);
    return bx;
}

// Canonical arrays of one box, in scratch order.
// compact transverse flux (one array per flux): FR FMX FMY FMZ FE Y0 Y1 Y2
// (no rho e), then the Godunov normal velocity and pressure
enum { TR = 0, TMX = 1, TMY = 2, TMZ = 3, TE = 4, TY = 5, TGU = 8, TGP = 9, NTF = 10 };
// final flux: the NVAR fluxes, then Godunov (u, p)
enum { XGU = NVAR, XGP = NVAR + 1, NXF = NVAR + 2 };
enum Slot {
    S_U,                 // Sborder copy (NVAR)
    S_Q,                 // compact primitives (NQC)
    S_SRCQ,              // compact primitive sources (NS)
    S_FLAT,              // flattening (1)
    S_QM0,               // qm per direction (NS)
    S_QP0 = S_QM0 + 3,   // qp per direction (NS)
    S_F1 = S_QP0 + 3,    // first fluxes per direction (NTF)
    S_F2 = S_F1 + 3,     // transdo-corrected fluxes, 6 (NTF)
    S_FLX = S_F2 + 6,    // final fluxes per direction (NXF)
    S_DIVU = S_FLX + 3,  // nodal divu (1)
    S_COUNT
};
__host__ __device__ constexpr int slot_comps(int s) {
    return s == S_U ? NVAR
         : s == S_Q ? NQC
         : s == S_SRCQ ? NS
         : s == S_FLAT ? 1
         : s < S_F1 ? NS
         : s < S_FLX ? NTF
         : s < S_DIVU ? NXF
         : 1;
}
__host__ __device__ constexpr int slot_off(int s) {
    int o = 0;
    for (int t = 0; t < s; ++t) o += slot_comps(t);
    return o;
}
// the slot of the d-flux corrected by direction t (t != d)
__host__ __device__ constexpr int f2_slot(int d, int t) {
    return S_F2 + 2 * d + (t > d ? t - 1 : t);
}

// One box of the level as the kernels see it.
struct BoxEnt {
    int id;         // box index in the level
    Box bx;         // valid box
    Canon L;        // layout of its canonical arrays
    double *base;   // its canonical arrays, in slot order
    Arr snew, sold; // level state (valid cells)
    double *abase;  // adjoints of the canonical arrays (same layout), or NULL
};

struct Batch {
    const BoxEnt *__restrict__ tab;
    const int *__restrict__ blk;  // block prefix of this stage kind (nbox + 1)
    int nbox;
    int first;                    // block id of this launch's first block
    const int *halt;              // graph replays: non-zero = skip (or NULL)
    // block shape: lx < 0 = the stage range in linear x-fastest order;
    // else a tile of 2^lx x 2^ly x 2^lz points per block
    int lx, ly, lz;
    // XCD chunk: 0, or K such that every K consecutive logical blocks run
    // on one XCD (see batch_box)
    int xk;
    // z-march: 0, or the number of consecutive z planes each thread walks
    // (the column kernels of batch_column)
    int zm;
};

// Number of blocks of one box's stage range sb under the block shape.
__host__ __device__ inline int64_t batch_blocks(const Box &sb, int block, int lx, int ly,
                                                int lz, int zm = 0) {
    if (zm > 0)
        return (((int64_t)sb.len(0) * sb.len(1) + block - 1) / block) *
               ((sb.len(2) + zm - 1) / zm);
    if (lx < 0) return (sb.npts() + block - 1) / block;
    return (int64_t)((sb.len(0) + (1 << lx) - 1) >> lx) * ((sb.len(1) + (1 << ly) - 1) >> ly) *
           ((sb.len(2) + (1 << lz) - 1) >> lz);
}

// Box of this block and the block's index within the box's stage range.
// The hardware hands workgroup w to XCD w % 8, so consecutive blocks (the
// neighbouring rows a stencil re-reads) land in eight different L2 caches.
// With B.xk = K the launch's blocks are permuted within groups of 8K: the
// workgroups of one XCD take K consecutive logical blocks, while all XCDs
// stay within the same window of 8K blocks (the same few planes).
__device__ __forceinline__ int batch_box(const Batch &B, int &lblock) {
    int w = (int)blockIdx.x;
    if (B.xk > 0) {
        const int g = 8 * B.xk;
        if (w < ((int)gridDim.x / g) * g) {
            const int r = w % g;
            w = (w - r) + (r & 7) * B.xk + (r >> 3);
        }
    }
    const int bid = w + B.first;
    int lo = 0, hi = B.nbox - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (B.blk[mid] <= bid) lo = mid;
        else hi = mid - 1;
    }
    lblock = bid - B.blk[lo];
    return lo;
}

// The point (i,j,k) of stage range sb that this thread covers as the
// lblk-th block of its box; false past the range's end.  Tiled blocks take
// their tiles x fastest; threads fill a tile x fastest, so a wavefront
// spans whole x rows of the tile.
__device__ __forceinline__ bool batch_point(const Batch &B, const Box &sb, int lblk, int &i,
                                            int &j, int &k) {
    if (B.lx < 0)
        return range_point(range_of(sb), (int64_t)lblk * blockDim.x + threadIdx.x, i, j, k);
    const int ntx = (sb.len(0) + (1 << B.lx) - 1) >> B.lx;
    const int nty = (sb.len(1) + (1 << B.ly) - 1) >> B.ly;
    const int bz = lblk / (ntx * nty);
    const int rem = lblk - bz * ntx * nty;
    const int by = rem / ntx;
    const int bx = rem - by * ntx;
    const int t = (int)threadIdx.x;
    i = sb.lo[0] + (bx << B.lx) + (t & ((1 << B.lx) - 1));
    j = sb.lo[1] + (by << B.ly) + ((t >> B.lx) & ((1 << B.ly) - 1));
    k = sb.lo[2] + (bz << B.lz) + (t >> (B.lx + B.ly));
    return i <= sb.hi[0] && j <= sb.hi[1] && k <= sb.hi[2];
}

// Column of a z-marching launch: the thread's (i, j) of the stage range's
// x-y plane and its run of planes k0..k1 (at most B.zm of them); false past
// the range's end.  A block's threads take consecutive points of the plane,
// x fastest, so each plane's loads stay row-coalesced.
__device__ __forceinline__ bool batch_column(const Batch &B, const Box &sb, int lblk, int &i,
                                             int &j, int &k0, int &k1) {
    const int n0 = sb.len(0);
    const int64_t np = (int64_t)n0 * sb.len(1);
    const int nbp = (int)((np + blockDim.x - 1) / blockDim.x);
    const int zc = lblk / nbp;
    const int64_t t = (int64_t)(lblk - zc * nbp) * blockDim.x + threadIdx.x;
    if (t >= np) return false;
    const int jj = (int)(t / n0);
    i = sb.lo[0] + (int)(t - (int64_t)jj * n0);
    j = sb.lo[1] + jj;
    k0 = sb.lo[2] + zc * B.zm;
    k1 = k0 + B.zm - 1 < sb.hi[2] ? k0 + B.zm - 1 : sb.hi[2];
    return true;
}

// Tiles of the LDS-staged CTU face kernels (k_ctu_x/y/z): kCtuX x kCtuY
// points of the x-y plane per block (two x rows per wavefront), each block
// walking `planes` z planes.  With merge (the x-face stage), a single point
// left over past the full x tiles joins the last full tile, which then is
// kCtuX + 1 points wide: the one face past a box's last cell would
// otherwise take a whole block.
constexpr int kCtuX = 32, kCtuY = 8, kCtuBlock = kCtuX * kCtuY;
__host__ __device__ inline int ctu_ntx(int n, bool merge) {
    return (merge && n > kCtuX && n % kCtuX == 1) ? n / kCtuX : (n + kCtuX - 1) / kCtuX;
}
__host__ __device__ inline int64_t ctu_blocks(const Box &sb, int planes, bool merge) {
    return (int64_t)ctu_ntx(sb.len(0), merge) * ((sb.len(1) + kCtuY - 1) / kCtuY) *
           ((sb.len(2) + planes - 1) / planes);
}
struct CtuTile {
    int i0, j0;  // first point of the tile
    int k0, k1;  // the block's planes
    int wx;      // points in x: kCtuX, fewer in a partial tile, kCtuX + 1 merged
};
__device__ __forceinline__ CtuTile ctu_tile(const Box &sb, int lblk, int planes, bool merge) {
    const int n0 = sb.len(0);
    const int ntx = ctu_ntx(n0, merge), nty = (sb.len(1) + kCtuY - 1) / kCtuY;
    const int bz = lblk / (ntx * nty);
    const int rem = lblk - bz * ntx * nty;
    const int by = rem / ntx, bx = rem - by * ntx;
    CtuTile t;
    t.i0 = sb.lo[0] + bx * kCtuX;
    t.j0 = sb.lo[1] + by * kCtuY;
    t.k0 = sb.lo[2] + bz * planes;
    t.k1 = t.k0 + planes - 1 < sb.hi[2] ? t.k0 + planes - 1 : sb.hi[2];
    const int rest = n0 - bx * kCtuX;
    t.wx = (bx == ntx - 1) ? rest : kCtuX;
    return t;
}

#define BATCH_COLUMN(KIND)                                                     \
    if (B.halt && *B.halt) return;                                             \
    int lblk_;                                                                 \
    const BoxEnt &E = B.tab[batch_box(B, lblk_)];                              \
    const Canon L = E.L;                                                       \
    int i, j, k0, k1;                                                          \
    if (!batch_column(B, stage_box((KIND), E.bx), lblk_, i, j, k0, k1)) return;

// First statements of every kernel: the block's box E (and its layout L)
// and the point (i,j,k) of the stage range this thread covers.
#define BATCH_POINT(KIND)                                                      \
    if (B.halt && *B.halt) return;                                             \
    int lblk_;                                                                 \
    const BoxEnt &E = B.tab[batch_box(B, lblk_)];                              \
    const Canon L = E.L;                                                       \
    int i, j, k;                                                               \
    if (!batch_point(B, stage_box((KIND), E.bx), lblk_, i, j, k)) return;

// Slot S of box E (S a compile-time constant, so the address stays scalar).
template <int S>
__device__ __forceinline__ double *slot_ptr(const BoxEnt &E) {
    constexpr size_t off = (size_t)slot_off(S);
    return E.base + off * E.L.cs;
}

// hydro_pdivu (Hydro_utils_K.H:175-202) and hydro_update (:239-257) of the
// valid cell (i,j,k) of box E on a zeroed dS/dt
// (CAMR_construct_hydro_source.cpp:14), then the Saxpy of
// CAMR_advance.cpp:141: v holds S_new on entry and S_new + dt dS/dt on
// return.
__device__ __forceinline__ void update_values(const BoxEnt &E, int i, int j, int k,
                                              const HydroConst &hc, double dt, double *v) {
    const Canon L = E.L;
    const double a0 = hc.dx[1] * hc.dx[2], a1 = hc.dx[0] * hc.dx[2], a2 = hc.dx[0] * hc.dx[1];
    const double vol = hc.dx[0] * hc.dx[1] * hc.dx[2];
    const CView f0 = cview(L, slot_ptr<S_FLX + 0>(E), NXF), f1 = cview(L, slot_ptr<S_FLX + 1>(E), NXF),
                f2 = cview(L, slot_ptr<S_FLX + 2>(E), NXF);
    const unsigned v0 = L.off(i, j, k), sx = 8u, sy = L.sy(), sz = L.sz();
    const double pdivu =
        0.5 *
        ((f0.ld(v0 + sx, XGP) + f0.ld(v0, XGP)) * (f0.ld(v0 + sx, XGU) * a0 - f0.ld(v0, XGU) * a0) +
         (f1.ld(v0 + sy, XGP) + f1.ld(v0, XGP)) * (f1.ld(v0 + sy, XGU) * a1 - f1.ld(v0, XGU) * a1) +
         (f2.ld(v0 + sz, XGP) + f2.ld(v0, XGP)) * (f2.ld(v0 + sz, XGU) * a2 - f2.ld(v0, XGU) * a2)) /
        vol;
    const double volinv = 1.0 / vol;
#pragma unroll
    for (int n = 0; n < NVAR; ++n) {
        double ds = 0.0;
        ds += (f0.ld(v0, n) - f0.ld(v0 + sx, n) + f1.ld(v0, n) - f1.ld(v0 + sy, n) +
               f2.ld(v0, n) - f2.ld(v0 + sz, n)) *
              volinv;
        if (n == UEINT) ds -= pdivu;
        v[n] += dt * ds;
    }
}

// Level-step kernels fusing the pointwise bookkeeping around the hydro
// source (level_fused.hip).
struct FillArgs {
    const DevFab *sold_tab;  // raw S_old of every box (valid cells)
    const int *owner;        // owning box of every domain cell (-1: none)
    Arr ext;                 // exterior data for cells no box owns
    int has_ext;
    DomInfo dom;
    HydroConst hc;
    const double *dtp;       // the step's dt (device)
    double g;
    int add_grav;
};
struct PostArgs {
    HydroConst hc;
    const double *dtp;
    double g;
    int add_grav;
    unsigned long long *dtmin;
    int defer_post;  // 1: leave post_timestep's clean_state to the caller
};

// Grid of one stage launch over a box set.
struct LaunchCfg {
    Batch B;
    unsigned grid;
    int block;
};
void launch_fill_fused(const LaunchCfg &c, const FillArgs &a, hipStream_t st, int *status);
void launch_post_fused(const LaunchCfg &c, const PostArgs &a, hipStream_t st, int *status);
// The update of k_update_c and the post pass in one kernel (S_new read and
// written once).
void launch_update_post_fused(const LaunchCfg &c, const HydroConst &hc, const double *dtp,
                              const PostArgs &a, hipStream_t st, int *status);

// Per-stage launch configuration (block size and shape per stage kind).
struct Tuning {
    int block[ST_COUNT];
    int order;  // 0: boxes in level order; 1: largest first
    // tile of each stage's blocks as log2 sizes (x, y, z) whose product is
    // the block size; tile[k][0] < 0: linear x-fastest blocks
    int tile[ST_COUNT][3];
    // XCD chunk of each stage (0: hardware order)
    int xcd[ST_COUNT];
    // z-march planes per thread (0: one point per thread); only the z-face
    // transverse stages (ST_T1 + 2, ST_T2 + 2) have marching kernels
    int zm[ST_COUNT];
    // 1: the transverse stages run as the LDS-staged CTU face kernels
    // (k_ctu_x/y/z, the transdo fluxes never stored); 0: the staged
    // k_trans1 / k_trans2 launches
    int ctu;
    int ctu_planes;  // z planes each CTU block walks
};

class FusedEngine {
  public:
    FusedEngine() = default;
    ~FusedEngine();
    FusedEngine(const FusedEngine &) = delete;
    FusedEngine &operator=(const FusedEngine &) = delete;
    static size_t scratch_doubles(const Box &bx);
    static bool fits(const Box &bx);
    // Canonical arrays for every box and the two box tables (state buffers
    // A/B as new/old and swapped).  ids: the level's index of each box (the
    // row of the level tables sb_tab / src_tab / dsdt_tab), default 0..n-1.
    int setup(const std::vector<Box> &boxes, const std::vector<Arr> &bufA,
              const std::vector<Arr> &bufB, const Tuning &t,
              const std::vector<int> &ids = std::vector<int>());
    int set_tuning(const Tuning &t);
    const Tuning &tuning() const { return tune_; }

    // Box set of one launch: b < 0 = all boxes (one launch per stage).
    // parity selects the table whose snew is buffer A (0) or B (1).
    // construct_hydro_source + the S_new Saxpy (snew += dt * dS/dt); floor
    // and energy checks into *status.  have_prim: fill() already wrote the
    // primitives and the Sborder copy; else they come from sb_tab (and the
    // primitive sources from src_tab when has_src).
    // With `post`, the update and the post pass (post()) run as one kernel;
    // with `no_update` the final update is left out (the fluxes are ready in
    // the scratch, S_new is not touched).
    int hydro(int b, int parity, bool have_prim, const DevFab *sb_tab, const DevFab *src_tab,
              int has_src, const HydroConst &hc, const double *dtp, hipStream_t st,
              int *status, Prof *prof, const PostArgs *post = nullptr,
              bool no_update = false);
    // ghost fill + clean + primitives + S_new = Sborder + dt * old source
    int fill(int b, int parity, const FillArgs &a, hipStream_t st, int *status, Prof *prof);
    // clean_state passes after the update + next dt estimate
    int post(int b, int parity, const PostArgs &a, hipStream_t st, int *status, Prof *prof);
    // after hydro(..., no_update = true): hydro_pdivu + hydro_update of
    // every box into dsdt_tab[id] (overwritten, the staged path's zero fill
    // and += order), for callers that apply dS/dt themselves
    int dsdt(int b, int parity, const DevFab *dsdt_tab, const HydroConst &hc, hipStream_t st,
             Prof *prof);
    int nbox() const { return (int)boxes_.size(); }
    // ghost width of the primitives the flattening stencil may read (q of
    // CAMR_construct_hydro_source.cpp, grow(bx, numGrow)): 4, or 6 on an EB
    // level, where points beyond the canonical grow(bx, 4) come from Sborder
    void set_qgrow(int g) { qgrow_ = g; }
    // adjoint arrays with the canonical layout of every box (allocated on
    // first use); total doubles of one scratch set
    int enable_adjoint();
    double *adjoint_base() const { return ascratch_.base(); }
    size_t scratch_total() const { return total_; }

    // linear = true: the stage's blocks without z-marching (kernels that
    // take one point per thread over a stage that marches when tuned)
    LaunchCfg cfg(int kind, int b, int parity, bool linear = false) const;
    // launches made while set read this word and skip their work when it is
    // non-zero (the halt flag of captured step graphs)
    void set_halt(const int *h) { halt_ = h; }
    // the final (adjusted, area-weighted) d-fluxes of box b as an Array4
    // view over the box's canonical box; components 0..NVAR-1 are the fluxes
    Arr flux_view(int b, int d) const {
        const Box cb = canon_box(boxes_[b]);
        return dense(scratch_.base() + off_[b] + (size_t)slot_off(S_FLX + d) * cb.npts(), cb);
    }

  private:
    template <typename K, typename... A>
    void launch(const char *name, Prof *prof, K kern, int kind, int b, int parity,
                hipStream_t st, A... args);
    // grid of the CTU face kernel of direction d
    LaunchCfg cfg_ctu(int d, int b, int parity) const;
    template <typename K, typename... A>
    void launch_ctu(const char *name, Prof *prof, K kern, int d, int b, int parity,
                    hipStream_t st, A... args);
    int build_tables();
    std::vector<Box> boxes_;
    std::vector<int> ids_;
    std::vector<Arr> bufA_, bufB_;
    std::vector<size_t> off_;      // scratch offset of each box (doubles)
    std::vector<int> order_;       // box order of the tables
    // (2 x ST_COUNT + 3) x (nbox + 1): as tuned, without z-march, CTU tiles
    std::vector<int> blk_host_;
    std::vector<int> pos_;         // table position of box b
    Tuning tune_{};
    BoxEnt *tab_[2] = {nullptr, nullptr};
    int *blk_dev_ = nullptr;
    const int *halt_ = nullptr;
    int qgrow_ = CAMR_NGROW;
    size_t total_ = 0;
    Scratch scratch_, ascratch_;
};

Tuning default_tuning();

// hydro_umdrv of one box (Hydro_umdrv.cpp:11-93, Godunov) on the fused
// engine's kernels, for callers that hold the reference's arrays: q / qaux
// (QVAR / NQAUX layout over exactly grow(bx, 4)), srcq, uin are converted
// into the canonical arrays, the fused stages run over a one-box table, and
// the fluxes go to flx[d] and the divergence is added to dsdt with the
// operation order of hydro_update.  Applies without area / volume arrays
// (uniform Cartesian grid) and without the method of lines.
struct UmdrvArgs;
bool fused_umdrv_applies(const UmdrvArgs &A, int do_mol);
size_t fused_umdrv_scratch_doubles(const Box &bx);
int fused_umdrv(const UmdrvArgs &A, double *scratch, hipStream_t st);

} // namespace camr
