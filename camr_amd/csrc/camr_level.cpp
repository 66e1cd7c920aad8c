// camr_level.cpp -- host side of the MI355X hydro advance: the C-ABI entry
// points of include/camr_hydro.h and the level driver that mirrors
// CAMR::CAMR_advance (Source/CAMR_advance.cpp:38-164),
// CAMR::construct_hydro_source (Hydro/CAMR_construct_hydro_source.cpp:9-267),
// CAMR::clean_state / expand_state (CAMR.cpp:1243-1262) and the dt control
// of CAMR::estTimeStep / computeNewDt / computeInitialDt (CAMR.cpp:474-676).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cfloat>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <sys/stat.h>
#include <mutex>
#include <vector>

#include "hydro_adjoint.h"
#include "hydro_eb_adjoint.h"
#include "hydro_fused.h"
#include "hydro_state.h"

namespace camr {

static thread_local char g_err[512] = "";

void set_error(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

#define CAMR_HIP(call)                                                        \
    do {                                                                      \
        hipError_t e_ = (call);                                               \
        if (e_ != hipSuccess) {                                               \
            set_error("HIP error %s at %s:%d", hipGetErrorString(e_),         \
                      __FILE__, __LINE__);                                    \
            return CAMR_ERR_HIP;                                              \
        }                                                                     \
    } while (0)

Scratch::~Scratch() {
    if (base_) (void)hipFree(base_);
}
bool Scratch::reserve(size_t bytes) {
    if (bytes <= cap_) return true;
    if (base_) (void)hipFree(base_);
    base_ = nullptr;
    cap_ = 0;
    if (hipMalloc(&base_, bytes) != hipSuccess) { base_ = nullptr; return false; }
    cap_ = bytes;
    return true;
}
double *Scratch::take(size_t n) {
    size_t bytes = ((n * sizeof(double)) + 255) & ~size_t(255);
    if (used_ + bytes > cap_) return nullptr;
    double *p = (double *)(base_ + used_);
    used_ += bytes;
    return p;
}

HydroConst make_const(const camr_hydro_params &p) {
    HydroConst h;
    h.ppm_type = p.ppm_type;
    h.plm_iorder = p.plm_iorder;
    h.use_pslope = p.use_pslope;
    h.use_flattening = p.use_flattening;
    h.trd = p.transverse_reset_density;
    h.allow_negative_energy = p.allow_negative_energy;
    h.allow_small_energy = p.allow_small_energy;
    h.update_E_from_e = p.dual_energy_update_E_from_e;
    h.small = p.small;
    h.small_dens = p.small_dens;
    h.small_pres = p.small_pres;
    h.small_temp = p.small_temp;
    h.smallu = p.smallu;
    h.small_num = p.small_num;
    h.difmag = p.difmag;
    h.eta1 = p.dual_energy_eta1;
    h.eta2 = p.dual_energy_eta2;
    for (int d = 0; d < 3; ++d) {
        h.bclo[d] = p.bclo[d];
        h.bchi[d] = p.bchi[d];
        h.domlo[d] = p.domlo[d];
        h.domhi[d] = p.domhi[d];
        h.dx[d] = p.dx[d];
    }
    h.eb = 0;
    return h;
}

static int check_params(const camr_hydro_params &p) {
    if (p.do_mol != 0 && p.do_mol != 1) {
        set_error("do_mol must be 0 (Godunov) or 1 (method of lines)");
        return CAMR_ERR_PARAM;
    }
    if (!p.do_mol && p.ppm_type != 0 && p.ppm_type != 1) {
        set_error("ppm_type must be 0 (PLM) or 1 (PPM)");  // CAMR.cpp:180-184
        return CAMR_ERR_PARAM;
    }
    if (p.plm_iorder != 1 && p.plm_iorder != 2 && p.plm_iorder != 4) {
        set_error("plm_iorder must be 1, 2 or 4");
        return CAMR_ERR_PARAM;
    }
    for (int d = 0; d < 3; ++d) {
        if (!(p.dx[d] > 0.0)) { set_error("dx must be positive"); return CAMR_ERR_PARAM; }
        if (p.bclo[d] < 0 || p.bclo[d] > 5 || p.bchi[d] < 0 || p.bchi[d] > 5) {
            set_error("invalid physical BC code");
            return CAMR_ERR_PARAM;
        }
    }
    return CAMR_OK;
}

// Write a double into device memory in stream order (no host staging).
static void put_double(double *dst, double v, hipStream_t s) {
    unsigned long long bits;
    memcpy(&bits, &v, sizeof(bits));
    launch_on(k_fill_u64, make_box1(), s, (unsigned long long *)dst, bits);
}

static bool covers(const camr_array4 &a, const Box &b, int nc) {
    return a.p != nullptr && box_of(a).contains(b) && a.ncomp >= nc;
}

// ============================================================ level driver
struct FluxReg;
class HydroLevel {
  public:
    explicit HydroLevel(const camr_level_desc &d) : desc_(d) {}
    ~HydroLevel() { release(); }

    int create();
    int upload(int b, const double *h);
    int download(int b, double *h);
    int state_array(int b, camr_array4 *out);
    int init_post();
    int initial_dt(double *dt);
    int new_dt(double dt_old, double *dt);
    int advance(double dt);
    int sync();
    void set_stream(hipStream_t s) { stream_ = s; }
    double time() const { return time_; }
    int set_time(double t) {
        int rc = resolve();
        if (rc) return rc;
        time_ = t;
        return CAMR_OK;
    }
    const camr_level_desc &desc() const { return desc_; }
    const std::vector<Box> &boxes() const { return box_; }
    bool is_eb() const { return ebflags_ != nullptr; }
    Prof &prof() { return prof_; }
    int set_eb(const int32_t *flags, const int32_t flo[3], const int32_t fhi[3]);
    int set_box_eb(int b, const double *vf, const double *const ap[3],
                   const double *const fc[3]);
    int box_kind(int b, int32_t *kind) const;
    int set_eb_merge(int on) {
        if (ebflags_) { set_error("set the merge of regular boxes before camr_level_set_eb"); return CAMR_ERR_PARAM; }
        eb_merge_ = on != 0;
        return CAMR_OK;
    }
    int eb_merge_info(int32_t *nvirtual, int32_t *nmerged) const {
        if (!nvirtual || !nmerged) { set_error("null argument"); return CAMR_ERR_PARAM; }
        *nvirtual = (int32_t)vbox_.size();
        *nmerged = nmerged_;
        return CAMR_OK;
    }
    int eb_force(double f[3]);

  private:
    int est_time_step(double *est);
    int advance_exact(double dt);
    int advance_fused(double dt);
    int resolve();
    int clean_state(std::vector<Arr> &S, bool launch_detect);
    int expand_state();
    int construct_hydro_source(double dt, bool add = true);
    int check_status();
    int status_error(int st);
    void release();
    Arr valid_view(double *p, const Box &b) const { return dense(p, b); }

    camr_level_desc desc_;
    HydroConst hc_{};
    DomInfo dom_{};
    std::vector<Box> box_;
    std::vector<double *> bufA_, bufB_, sb_, dsdt_, srchyd_;
    std::vector<Arr> dsv_;  // dS/dt of every box (a view into its virtual box's on an EB level)
    bool a_is_new_ = true;
    std::vector<Arr> snew_, sold_, sbv_;
    DevFab *tab_old_ = nullptr, *tab_sb_ = nullptr, *tab_src_ = nullptr;
    int *owner_ = nullptr;
    int *status_ = nullptr;              // device: error bits | floor flag
    int *redo_ = nullptr;                // device: the staged tail must be repeated pass by pass
    unsigned long long *dtmin_ = nullptr;  // device: bit pattern of min dt
    int *h_status_ = nullptr;            // pinned host mirrors
    unsigned long long *h_dtmin_ = nullptr;
    double *q_ = nullptr, *qaux_ = nullptr, *srcq_ = nullptr, *flx_[3] = {};
    Scratch scratch_;
    hipStream_t stream_ = nullptr;
    double time_ = 0.0;
    bool grav_src_values_ = false;
    Prof prof_;
    bool fused_ = true;
    FusedEngine engine_;
    // embedded boundary (AMREX_USE_EB build): ghost width 6, flags over the
    // domain grown by 7, per-box path (0 regular, 1 cut cells, 2 covered)
    int ng_ = CAMR_NGROW;
    // staged engine: per-box work spread over worker streams (default
    // priority), each with its own q / qaux / srcq and scratch; events join
    // them where boxes read each other's data
    static constexpr int kWorkers = 16;  // most worker streams
    int nworkers() const { return desc_.worker_streams > 0 ? desc_.worker_streams : 4; }
    int nws_ = 0;
    hipStream_t wst_[kWorkers] = {};
    hipEvent_t wev_[kWorkers] = {}, mev_ = nullptr;
    size_t qstride_ = 0, sstride_ = 0;
    // worker of box b: b mod nws_, or wmap_[b] when set (an EB level keeps its
    // cut-cell boxes off worker 0, which runs the batch of regular boxes)
    std::vector<int> wmap_;
    int wslot(size_t b) const {
        return nws_ ? (wmap_.empty() ? (int)(b % nws_) : wmap_[b]) : 0;
    }
    hipStream_t ws(size_t b) const { return nws_ ? wst_[wslot(b)] : stream_; }
    bool in_ws_ = false;  // inside a step spread over the worker streams
    hipStream_t cur_ws(size_t b) const { return in_ws_ ? ws(b) : stream_; }
    size_t fstride_ = 0;
    int fork();
    int join();
    int setup_workers(size_t max_g, size_t sc);
    int32_t *ebflags_ = nullptr;
    EbFlags ebf_{};
    std::vector<int> kind_;
    // EB level: the all-regular boxes (kind 0) as one box batch of the
    // fused engine's kernels (hydro_umdrv of CAMR_construct_hydro_source.cpp:
    // 150-151 for each, in batched launches); the cut-cell boxes keep the
    // per-box EB path
    FusedEngine reg_engine_;
    bool eb_batched_ = false;
    // regular boxes of equal shape that tile a box in blocks of up to
    // 4x4x4 run in that batch as the one virtual box they tile (fewer ghost
    // cells): its Sborder over grow(v, 6) and its dS/dt are one array each,
    // which their rows of tab_sb_ / tab_dsdt_ view; expand_state fills the
    // virtual box's Sborder as one row of tab_exp_ (exp_box_: the rows'
    // boxes).  tab_sbx_ / tab_dsdtx_: the level's rows, then one per
    // virtual box
    bool eb_merge_ = true;
    std::vector<Box> vbox_;
    int nmerged_ = 0;
    int *vg_ids_ = nullptr;
    std::vector<int> vof_;       // virtual box of each level box, or -1
    std::vector<VGather> vgh_;   // each virtual box's block and Sborder
    std::vector<Arr> vds_;       // dS/dt of each virtual box
    VGather *vga_dev_ = nullptr; // the adjoint's: sb = dL/dSborder of each virtual box
    DevFab *tab_sbx_ = nullptr, *tab_dsdtx_ = nullptr;
    DevFab *tab_exp_ = nullptr;
    std::vector<Box> exp_box_;
    int eb_regular_batch();
    // points of the largest grow(v, 6) (0: no virtual box)
    int64_t vbox_points() const {
        int64_t n = 0;
        for (const Box &V : vbox_) n = std::max<int64_t>(n, grow(V, ng_).npts());
        return n;
    }
    std::vector<EbGeom> ebg_;
    std::vector<double *> ebbuf_;
    // the cut-cell boxes (kind 1) as one table of the batched EB path, each
    // with its own q / qaux / srcq, fluxes and scratch in cut_buf_ (freed
    // with ebbuf_)
    std::vector<int> cut_ids_;
    std::vector<size_t> cut_off_;
    std::vector<EbBox> cut_htab_;
    EbBox *cut_tab_ = nullptr;  // device copy of cut_htab_
    double *cut_buf_ = nullptr;
    CutCell *cut_cells_ = nullptr;  // the cut cells of grow(bx, 2) of the table's boxes
    int ncut_cells_ = 0;
    std::vector<CutCell> cut_cells_h_;  // host copy (entries: cut-table positions)
    CutCell *eadj_cells_ = nullptr;     // the same for the adjoint's table
    int neadj_cells_ = 0;
    int cut_entry(int c);
    bool pending_ = false;   // a fused step awaits resolve()
    bool needs_ext_ = false; // some ghost cell has no owning box
    bool dt_ready_ = false;  // dtmin_ holds the CFL estimate of S_new
    double last_dt_ = 0.0;
    DevFab *tab_a_ = nullptr, *tab_b_ = nullptr;  // raw-state tables of buffers A / B
    DevFab *tab_dsdt_ = nullptr;                  // staged engine: dS/dt of every box
    int advance_staged(double dt);
    // the staged step leaves the CFL estimate of S_new in dtmin_ (its
    // one-pass tail with post_timestep's clean)
    bool staged_dt() const { return !defer_post_; }
    void staged_tail(const DevFab *tnew, const DevFab *told, int mol);
    // cleaned: src went through clean_all (kb_expand_cleaned, one pass)
    void expand_all(const DevFab *src, bool cleaned = false);
    int hydro_on_workers(double dt);
    // staged engine: the step captured once per buffer parity, replayed
    // with the step's dt in dt_dev_
    // [buffer parity][post_timestep clean deferred]
    hipGraph_t sgraph_[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};
    hipGraphExec_t sgexec_[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};
    bool use_staged_graph() const { return desc_.graph != 0 && !fused_ && !grav_src_values_; }
    int replay_staged(double dt);
    int clean_all(const DevFab *tab, bool launch_detect);
    template <typename K, typename... A>
    void launch_all(K kern, int g, A... args);
    double *dt_dev_ = nullptr;  // the step's dt, read by the kernels
    double *ext_ = nullptr;     // exterior data (whole domain), when set
    double *cstage_ = nullptr;  // coarse values staged for exterior_from
    size_t cstage_n_ = 0;
    // the domain cells, in domain order, that the boxes' ghost regions map
    // to and no box owns (x, y, z triples; built on the first ghost fill)
    int *shell_cells_ = nullptr;
    int64_t shell_n_ = -1;
    // the coarse cells (grown coarse domain) their interpolation reads, for
    // the refinement ratio shell_ratio_
    int *cstage_cells_ = nullptr;
    int64_t cstage_cells_n_ = 0;
    int shell_ratio_ = 0;
    int build_shell_cells(int ratio, const Box &cdom);
    void drop_shell_cells();
    Arr extv_{};
    bool batched_ = true;       // one launch per stage over all boxes
    int parity() const { return a_is_new_ ? 0 : 1; }
    void set_dt_dev(double dt);

    // captured level-step graphs (one per buffer parity)
    StepCtl *ctl_ = nullptr;
    double *ctl_dts_ = nullptr;
    int ctl_cap_ = 0;
    hipStream_t gstream_ = nullptr;
    hipGraph_t graph_[2] = {nullptr, nullptr};
    hipGraphExec_t gexec_[2] = {nullptr, nullptr};
    int capture_graphs();
    void release_graphs();

    // adjoint: state adjoints (two sets per box), device checkpoints
    std::vector<double *> abar_[2];
    DevFab *tab_abar_[2] = {nullptr, nullptr};
    int acur_ = 0;                      // set holding the current adjoint
    std::vector<double *> ckpt_, blk_;  // whole-level state copies
    std::vector<double> adj_dts_;
    int adj_every_ = 0;
    bool adj_tail_ready_ = false;  // blk_ holds the last block's states from the forward run
    double adj_time0_ = 0.0;
    size_t state_doubles_ = 0;
    int64_t *state_off_ = nullptr;  // device: each box's offset in a flat state copy
    // One launch over every box's state (grid.y = box), a grid-stride x.
    template <typename K, typename... A>
    void launch_flat(K kern, A... args) {
        int64_t n = 1;
        for (const Box &B : box_) n = std::max<int64_t>(n, (int64_t)NVAR * B.npts());
        const unsigned gx = (unsigned)std::min<int64_t>((n + kBlock - 1) / kBlock, 2048);
        hipLaunchKernelGGL(kern, dim3(gx, (unsigned)box_.size()), dim3(kBlock), 0, stream_, args...);
    }
    // EB level: the step's adjoint through the cut-cell stages, for every
    // box that is not covered (the all-regular boxes with regular geometry,
    // which the cut-cell stages then compute as the regular path does)
    std::vector<int> eadj_ids_;
    std::vector<EbBox> eadj_fh_;
    EbBox *eadj_ftab_ = nullptr;
    std::vector<EbAdjEnt> eadj_h_;
    EbAdjEnt *eadj_tab_ = nullptr;
    double *eadj_fbuf_ = nullptr, *eadj_abuf_ = nullptr;  // forward / adjoint arrays
    size_t eadj_an_ = 0;                                   // doubles of eadj_abuf_
    // the parts of eadj_abuf_ cleared before each adjoint step: all but the
    // pair and final face-state adjoints, which their Riemann stages set
    // (range starts and lengths in doubles)
    std::vector<size_t> eadj_zs_, eadj_zl_;
    DevFab *tab_ub_ = nullptr, *tab_sbb_ = nullptr;
    EbGeom *geom_dev_ = nullptr;
    int *kinds_dev_ = nullptr;
    EbGeom unit_geom_{};
    int eb_adjoint_setup();
    int adjoint_step_eb(double dt, bool check);
    int adjoint_scope() const;
    int adjoint_alloc(int nck, int every);
    int save_state(double *dst);
    int load_state(const double *src);
    int adjoint_step(double dt, bool check = true);
    int adjoint_flags();

  public:
    int adjoint_forward(int nsteps, const double *dts, int every);
    int eb_force_seed(const double w[3]);
    int adjoint_backward();
    int adj_upload(int b, const double *h);
    int adj_download(int b, double *h);
    int run(int nsteps, double *dts);
    int run_graph(int nsteps, double *dts);
    bool use_graph() const {
        return desc_.graph != 0 && fused_ && batched_ && !grav_src_values_;
    }
    int set_exterior(const double *h);
    int exterior_from(HydroLevel &coarse, int ratio, int tmode = 0, double ta = 0.0,
                      double tb = 0.0, bool ghosts_only = false);
    int set_tuning(const Tuning &t);
    int get_tuning(Tuning *t) const;
    bool fused() const { return fused_; }

  private:
    // coarse-fine synchronisation: the register this level feeds as the
    // coarse level (CrseAdd) and as the fine level (FineAdd)
    FluxReg *reg_crse_ = nullptr, *reg_fine_ = nullptr;
    bool defer_post_ = false;  // leave post_timestep's clean_state to amr_step
    void reg_add(size_t b, const Arr *flx, hipStream_t s);
    // the fused engine keeps the step's fluxes in its canonical arrays: add
    // them once the step is final (after resolve)
    int reg_add_fused();
    friend struct FluxReg;
    friend int flux_register_create(HydroLevel &C, HydroLevel &F, int ratio, FluxReg **out);
    friend int average_down(HydroLevel &C, HydroLevel &F, int ratio);
    friend int amr_step(HydroLevel &C, HydroLevel &F, FluxReg &R, double dtc);
    friend int amr_advance_level(const std::vector<HydroLevel *> &L, const std::vector<FluxReg *> &R,
                                 size_t l, double dt);
    friend int amr_advance(const std::vector<HydroLevel *> &L, const std::vector<FluxReg *> &R,
                           double dt0);
};

// The flux register between a coarse and a fine level (see the
// coarse-fine synchronisation section below).
struct FluxReg {
    HydroLevel *C = nullptr, *F = nullptr;
    int ratio = 2;
    uint8_t *flag = nullptr;  // device, per coarse domain cell
    // crse / cfp per coarse box, views into one pool (reset in one launch)
    std::vector<double *> crse, cfp;
    double *pool = nullptr;
    int64_t pool_n = 0;
    DevFab *tab_crse = nullptr, *tab_cfp = nullptr;
    double dx1c = 0.0, dx1f = 0.0;
    // per fine box, bit 2d+side: its slab on that side holds boundary cells
    std::vector<int> slabs;
    // one-launch add records for fused-engine levels (their fluxes live in
    // fixed canonical arrays): CrseAdd per coarse box, FineAdd per slab
    RegCrseRec *crse_recs = nullptr;
    RegFineRec *fine_recs = nullptr;
    int ncrse_recs = 0, nfine_recs = 0;
    int64_t crse_pts = 0, fine_pts = 0;  // largest record range
    ~FluxReg() {
        // captured level steps hold the register's launches
        if (C && C->reg_crse_ == this) { C->release_graphs(); C->reg_crse_ = nullptr; }
        if (F && F->reg_fine_ == this) { F->release_graphs(); F->reg_fine_ = nullptr; }
        auto fr = [](void *p) { if (p) (void)hipFree(p); };
        fr(pool);
        fr(flag);
        fr(tab_crse);
        fr(tab_cfp);
        fr(crse_recs);
        fr(fine_recs);
    }
};

void HydroLevel::set_dt_dev(double dt) {
    unsigned long long bits;
    memcpy(&bits, &dt, sizeof(bits));
    launch_on(k_fill_u64, make_box1(), stream_, (unsigned long long *)dt_dev_, bits);
}

// Exterior data: values of the cells of the domain that no box owns, used
// as the FillPatch result for ghost cells there (coarse-fine ghosts of a
// refined level).  `h` holds NVAR components over the whole domain.
int HydroLevel::set_exterior(const double *h) {
    if (!h) { set_error("null exterior data"); return CAMR_ERR_PARAM; }
    int rc = resolve();
    if (rc) return rc;
    const Box dom = make_box(desc_.hp.domlo, desc_.hp.domhi);
    if (!ext_) {
        release_graphs();  // captured steps were recorded without exterior data
        CAMR_HIP(hipMalloc(&ext_, NVAR * dom.npts() * sizeof(double)));
    }
    CAMR_HIP(hipMemcpy(ext_, h, NVAR * dom.npts() * sizeof(double), hipMemcpyHostToDevice));
    extv_ = dense(ext_, dom);
    return CAMR_OK;
}

// Exterior data from the coarser level: FillPatch's coarse-fine
// interpolation (k_ext_from_coarse) at every cell of this level's domain,
// from the coarse level's current S_new.
int HydroLevel::exterior_from(HydroLevel &C, int ratio, int tmode, double ta, double tb,
                              bool ghosts_only) {
    if (ratio < 2) { set_error("refinement ratio must be at least 2"); return CAMR_ERR_PARAM; }
    // EB builds interpolate with eb_cell_cons_interp (CAMR_setup.cpp:139-143),
    // which weighs by volume fractions and skips covered cells: not built
    if (ebflags_ || C.ebflags_) {
        set_error("coarse-fine interpolation with embedded boundaries is not built");
        return CAMR_ERR_PARAM;
    }
    for (int d = 0; d < 3; ++d)
        if (desc_.hp.domlo[d] != C.desc_.hp.domlo[d] * ratio ||
            desc_.hp.domhi[d] + 1 != (C.desc_.hp.domhi[d] + 1) * ratio) {
            set_error("the fine domain is not the coarse domain refined by %d", ratio);
            return CAMR_ERR_SHAPE;
        }
    int rc = C.resolve();
    if (rc) return rc;
    CAMR_HIP(hipStreamSynchronize(C.stream_));
    rc = resolve();
    if (rc) return rc;
    const Box dom = make_box(desc_.hp.domlo, desc_.hp.domhi);
    if (!ext_) {
        release_graphs();  // captured steps were recorded without exterior data
        CAMR_HIP(hipMalloc(&ext_, NVAR * dom.npts() * sizeof(double)));
    }
    extv_ = dense(ext_, dom);
    const DevFab *ct = C.a_is_new_ ? C.tab_a_ : C.tab_b_;
    const DevFab *co = C.a_is_new_ ? C.tab_b_ : C.tab_a_;
    // pass 1: the coarse values the interpolation reads (source cell, fill
    // time, reflection sign) staged densely over the coarse domain grown by 1
    const Box cst = grow(make_box(C.desc_.hp.domlo, C.desc_.hp.domhi), 1);
    if (!cstage_ || cstage_n_ < NVAR * cst.npts()) {
        if (cstage_) (void)hipFree(cstage_);
        cstage_ = nullptr;
        CAMR_HIP(hipMalloc(&cstage_, NVAR * cst.npts() * sizeof(double)));
        cstage_n_ = NVAR * cst.npts();
    }
    const Arr stv = dense(cstage_, cst);
    if (ghosts_only) {
        // just the cells a fill reads from the exterior: the domain images
        // of the boxes' ghost regions that no box owns, and the coarse
        // cells their interpolation reads
        if (shell_n_ < 0 || shell_ratio_ != ratio) {
            rc = build_shell_cells(ratio, make_box(C.desc_.hp.domlo, C.desc_.hp.domhi));
            if (rc) return rc;
        }
        if (cstage_cells_n_ > 0)
            hipLaunchKernelGGL(k_coarse_stage_cells, dim3(grid_for(cstage_cells_n_), NVAR),
                               dim3(kBlock), 0, stream_, (const int *)cstage_cells_,
                               cstage_cells_n_, stv, ct, (const int *)C.owner_, C.dom_, co, tmode,
                               ta, tb);
        // pass 2: one thread per fine cell and component (grid.y = NVAR)
        if (shell_n_ > 0)
            hipLaunchKernelGGL(k_ext_cells, dim3(grid_for(shell_n_), NVAR), dim3(kBlock), 0,
                               stream_, (const int *)shell_cells_, shell_n_, extv_, stv, C.dom_,
                               ratio, dom.lo[0], dom.lo[1], dom.lo[2]);
    } else {
        {
            const Range r = range_of(cst);
            hipLaunchKernelGGL(k_coarse_stage, dim3(grid_for(r.n), NVAR), dim3(kBlock), 0, stream_,
                               r, stv, ct, (const int *)C.owner_, C.dom_, co, tmode, ta, tb);
        }
        const Range r = range_of(dom);
        hipLaunchKernelGGL(k_ext_from_coarse, dim3(grid_for(r.n), NVAR), dim3(kBlock), 0, stream_,
                           r, extv_, stv, C.dom_, ratio, dom.lo[0], dom.lo[1], dom.lo[2],
                           (const int *)nullptr, dom_);
    }
    CAMR_HIP(hipGetLastError());
    // the coarse level may step next: finish reading its state first
    CAMR_HIP(hipStreamSynchronize(stream_));
    return CAMR_OK;
}

void HydroLevel::drop_shell_cells() {
    if (shell_cells_) (void)hipFree(shell_cells_);
    if (cstage_cells_) (void)hipFree(cstage_cells_);
    shell_cells_ = cstage_cells_ = nullptr;
    shell_n_ = -1;
    cstage_cells_n_ = 0;
    shell_ratio_ = 0;
}

int HydroLevel::build_shell_cells(int ratio, const Box &cdom) {
    drop_shell_cells();
    const Box dom = make_box(desc_.hp.domlo, desc_.hp.domhi);
    const int64_t n0 = dom.len(0), n1 = dom.len(1), n2 = dom.len(2);
    std::vector<int> own((size_t)dom.npts());
    CAMR_HIP(hipMemcpy(own.data(), owner_, own.size() * sizeof(int), hipMemcpyDeviceToHost));
    std::vector<uint8_t> mark(own.size(), 0);
    const int64_t len[3] = {n0, n1, n2};
    for (const Box &B0 : box_) {
        const Box G = grow(B0, ng_);
        for (int k = G.lo[2]; k <= G.hi[2]; ++k)
            for (int j = G.lo[1]; j <= G.hi[1]; ++j)
                for (int i = G.lo[0]; i <= G.hi[0]; ++i) {
                    const int c[3] = {i, j, k};
                    int64_t m[3];
                    bool in = true;
                    for (int d = 0; d < 3 && in; ++d) {
                        int64_t x = c[d] - dom.lo[d];
                        if (x < 0 || x >= len[d]) {
                            if (!desc_.is_periodic[d]) in = false;
                            else x = (x % len[d] + len[d]) % len[d];
                        }
                        m[d] = x;
                    }
                    if (!in) continue;
                    const int64_t at = m[0] + m[1] * n0 + m[2] * n0 * n1;
                    if (own[(size_t)at] < 0) mark[(size_t)at] = 1;
                }
    }
    // the coarse parents of the listed cells and their 26 neighbours, over
    // the coarse domain grown by one (the staged array's box)
    const Box cst = grow(cdom, 1);
    const int64_t c0 = cst.len(0), c1 = cst.len(1), c2 = cst.len(2);
    std::vector<uint8_t> cmark((size_t)cst.npts(), 0);
    std::vector<int> cells;
    for (int64_t k = 0; k < n2; ++k)
        for (int64_t j = 0; j < n1; ++j)
            for (int64_t i = 0; i < n0; ++i)
                if (mark[(size_t)(i + j * n0 + k * n0 * n1)]) {
                    cells.push_back((int)i + dom.lo[0]);
                    cells.push_back((int)j + dom.lo[1]);
                    cells.push_back((int)k + dom.lo[2]);
                    // parent in the staged box's index space (cst.lo = cdom.lo - 1)
                    const int64_t pi = i / ratio + 1, pj = j / ratio + 1, pk = k / ratio + 1;
                    for (int64_t dk = -1; dk <= 1; ++dk)
                        for (int64_t dj = -1; dj <= 1; ++dj)
                            for (int64_t di = -1; di <= 1; ++di)
                                cmark[(size_t)((pi + di) + (pj + dj) * c0 + (pk + dk) * c0 * c1)] = 1;
                }
    shell_n_ = (int64_t)(cells.size() / 3);
    if (shell_n_ > 0) {
        CAMR_HIP(hipMalloc(&shell_cells_, cells.size() * sizeof(int)));
        CAMR_HIP(hipMemcpy(shell_cells_, cells.data(), cells.size() * sizeof(int),
                           hipMemcpyHostToDevice));
    }
    std::vector<int> ccells;
    for (int64_t k = 0; k < c2; ++k)
        for (int64_t j = 0; j < c1; ++j)
            for (int64_t i = 0; i < c0; ++i)
                if (cmark[(size_t)(i + j * c0 + k * c0 * c1)]) {
                    ccells.push_back((int)i + cst.lo[0]);
                    ccells.push_back((int)j + cst.lo[1]);
                    ccells.push_back((int)k + cst.lo[2]);
                }
    cstage_cells_n_ = (int64_t)(ccells.size() / 3);
    if (cstage_cells_n_ > 0) {
        CAMR_HIP(hipMalloc(&cstage_cells_, ccells.size() * sizeof(int)));
        CAMR_HIP(hipMemcpy(cstage_cells_, ccells.data(), ccells.size() * sizeof(int),
                           hipMemcpyHostToDevice));
    }
    shell_ratio_ = ratio;
    return CAMR_OK;
}

// The launch shapes of the level's fused kernels: the fused engine, or on
// an EB level the batch of its regular boxes.
int HydroLevel::set_tuning(const Tuning &t) {
    FusedEngine *e = fused_ ? &engine_ : (eb_batched_ ? &reg_engine_ : nullptr);
    if (!e) { set_error("tuning applies to the fused engine or an EB level's regular boxes"); return CAMR_ERR_PARAM; }
    int rc = resolve();
    if (rc) return rc;
    CAMR_HIP(hipStreamSynchronize(stream_));
    release_graphs();
    return e->set_tuning(t);
}
int HydroLevel::get_tuning(Tuning *t) const {
    const FusedEngine *e = fused_ ? &engine_ : (eb_batched_ ? &reg_engine_ : nullptr);
    if (!e) { set_error("tuning applies to the fused engine or an EB level's regular boxes"); return CAMR_ERR_PARAM; }
    *t = e->tuning();
    return CAMR_OK;
}

void HydroLevel::release() {
    // a register outliving its levels no longer reaches them
    if (reg_crse_) reg_crse_->C = nullptr;
    if (reg_fine_) reg_fine_->F = nullptr;
    reg_crse_ = reg_fine_ = nullptr;
    auto fr = [](void *p) { if (p) (void)hipFree(p); };
    fr(cstage_);
    cstage_ = nullptr;
    drop_shell_cells();
    for (auto p : bufA_) fr(p);
    for (auto p : bufB_) fr(p);
    for (auto p : sb_) fr(p);
    for (auto p : dsdt_) fr(p);
    for (auto p : srchyd_) fr(p);
    bufA_.clear(); bufB_.clear(); sb_.clear(); dsdt_.clear(); srchyd_.clear(); dsv_.clear();
    vbox_.clear();
    nmerged_ = 0;
    vg_ids_ = nullptr;  // freed with ebbuf_
    vga_dev_ = nullptr;
    vof_.clear();
    vgh_.clear();
    vds_.clear();
    tab_sbx_ = tab_dsdtx_ = nullptr;
    tab_exp_ = nullptr;
    exp_box_.clear();
    fr(tab_old_); fr(tab_sb_); fr(tab_src_); fr(tab_a_); fr(tab_b_); fr(tab_dsdt_);
    tab_dsdt_ = nullptr;
    tab_a_ = tab_b_ = nullptr; fr(owner_); fr(status_); fr(redo_); fr(dtmin_);
    redo_ = nullptr;
    fr(dt_dev_); fr(ext_); dt_dev_ = nullptr; ext_ = nullptr;
    release_graphs();
    fr(ctl_); fr(ctl_dts_); ctl_ = nullptr; ctl_dts_ = nullptr; ctl_cap_ = 0;
    for (int q = 0; q < 2; ++q) {
        for (auto p : abar_[q]) fr(p);
        abar_[q].clear();
        fr(tab_abar_[q]);
        tab_abar_[q] = nullptr;
    }
    for (auto p : ckpt_) fr(p);
    for (auto p : blk_) fr(p);
    fr(state_off_);
    state_off_ = nullptr;
    ckpt_.clear();
    blk_.clear();
    if (gstream_) (void)hipStreamDestroy(gstream_);
    gstream_ = nullptr;
    fr(q_); fr(qaux_); fr(srcq_);
    for (auto &f : flx_) { fr(f); f = nullptr; }
    fr(ebflags_); ebflags_ = nullptr;
    for (int w = 0; w < kWorkers; ++w) {
        if (wst_[w]) (void)hipStreamDestroy(wst_[w]);
        if (wev_[w]) (void)hipEventDestroy(wev_[w]);
        wst_[w] = nullptr;
        wev_[w] = nullptr;
    }
    if (mev_) (void)hipEventDestroy(mev_);
    mev_ = nullptr;
    nws_ = 0;
    for (auto p : ebbuf_) fr(p);
    ebbuf_.clear();
    if (h_status_) (void)hipHostFree(h_status_);
    if (h_dtmin_) (void)hipHostFree(h_dtmin_);
    tab_old_ = tab_sb_ = tab_src_ = nullptr;
    owner_ = status_ = nullptr;
    dtmin_ = nullptr;
    h_status_ = nullptr;
    h_dtmin_ = nullptr;
    q_ = qaux_ = srcq_ = nullptr;
}

int HydroLevel::create() {
    int rc = check_params(desc_.hp);
    if (rc) return rc;
    if (desc_.nbox <= 0 || !desc_.box_lohi) { set_error("no boxes"); return CAMR_ERR_PARAM; }
    if (!(desc_.cfl > 0.0 && desc_.cfl <= 1.0)) {
        set_error("Invalid CFL factor; must be between zero and one.");  // CAMR.cpp:173-174
        return CAMR_ERR_PARAM;
    }
    if (desc_.worker_streams < 0 || desc_.worker_streams > kWorkers) {
        set_error("worker_streams must be 0 (default 4) .. %d", kWorkers);
        return CAMR_ERR_PARAM;
    }
    if (desc_.hp.do_mol && desc_.cfl > 0.3 && desc_.fixed_dt <= 0.0) {
        set_error("Invalid CFL factor; must be <= 0.3 when using MOL hydro");  // CAMR.cpp:175-177
        return CAMR_ERR_PARAM;
    }
    if (desc_.hp.do_mol && desc_.engine == 2) {
        set_error("the method of lines runs on the staged engine");
        return CAMR_ERR_PARAM;
    }
    hc_ = make_const(desc_.hp);
    for (int d = 0; d < 3; ++d) {
        dom_.domlo[d] = desc_.hp.domlo[d];
        dom_.domhi[d] = desc_.hp.domhi[d];
        dom_.periodic[d] = desc_.is_periodic[d];
        dom_.bclo[d] = desc_.hp.bclo[d];
        dom_.bchi[d] = desc_.hp.bchi[d];
        if (desc_.is_periodic[d] && (desc_.hp.bclo[d] != 0 || desc_.hp.bchi[d] != 0)) {
            set_error("periodic direction needs Interior BCs");  // CAMR.cpp:129-143
            return CAMR_ERR_PARAM;
        }
        if (!desc_.is_periodic[d] && (desc_.hp.bclo[d] == 0 || desc_.hp.bchi[d] == 0)) {
            set_error("Interior BC on a non-periodic direction");
            return CAMR_ERR_PARAM;
        }
    }
    const Box dom = make_box(desc_.hp.domlo, desc_.hp.domhi);
    const int nb = desc_.nbox;
    size_t max_g4 = 0, max_valid = 0, max_nodes = 0;
    Box maxbox{};
    for (int b = 0; b < nb; ++b) {
        Box B = make_box(desc_.box_lohi + 6 * b, desc_.box_lohi + 6 * b + 3);
        if (!B.ok() || !dom.contains(B)) { set_error("box %d outside the domain", b); return CAMR_ERR_PARAM; }
        box_.push_back(B);
        max_g4 = std::max<size_t>(max_g4, grow(B, CAMR_NGROW).npts());
        max_valid = std::max<size_t>(max_valid, B.npts());
        for (int d = 0; d < 3; ++d) max_nodes = std::max<size_t>(max_nodes, nodes(B, d).npts());
        if (B.npts() > maxbox.npts()) maxbox = B;
    }
    grav_src_values_ = desc_.add_grav_src && desc_.const_grav != 0.0;
    auto alloc = [&](double **p, size_t n) -> int {
        if (hipMalloc(p, std::max<size_t>(n, 1) * sizeof(double)) != hipSuccess) {
            set_error("device allocation of %zu doubles failed", n);
            return CAMR_ERR_ALLOC;
        }
        return CAMR_OK;
    };
    for (int b = 0; b < nb; ++b) {
        double *p;
        const Box B = box_[b];
        if ((rc = alloc(&p, NVAR * B.npts()))) return rc;
        bufA_.push_back(p);
        if ((rc = alloc(&p, NVAR * B.npts()))) return rc;
        bufB_.push_back(p);
        if ((rc = alloc(&p, NVAR * grow(B, CAMR_NGROW).npts()))) return rc;
        sb_.push_back(p);
        if ((rc = alloc(&p, NVAR * B.npts()))) return rc;
        dsdt_.push_back(p);
        dsv_.push_back(dense(p, B));
        if (grav_src_values_) {
            if ((rc = alloc(&p, NVAR * grow(B, CAMR_NGROW).npts()))) return rc;
            srchyd_.push_back(p);
        }
        sbv_.push_back(dense(sb_[b], grow(B, CAMR_NGROW)));
    }
    snew_.resize(nb);
    sold_.resize(nb);
    for (int b = 0; b < nb; ++b) {
        snew_[b] = dense(bufA_[b], box_[b]);
        sold_[b] = dense(bufB_[b], box_[b]);
    }
    // owner map over the domain
    const size_t ncell = dom.npts();
    std::vector<int> own(ncell, -1);
    const int64_t n0 = dom.len(0), n1 = dom.len(1);
    for (int b = 0; b < nb; ++b) {
        const Box B = box_[b];
        for (int k = B.lo[2]; k <= B.hi[2]; ++k)
            for (int j = B.lo[1]; j <= B.hi[1]; ++j)
                for (int i = B.lo[0]; i <= B.hi[0]; ++i) {
                    int &o = own[(i - dom.lo[0]) + (j - dom.lo[1]) * n0 + (int64_t)(k - dom.lo[2]) * n0 * n1];
                    if (o >= 0) { set_error("boxes %d and %d overlap", o, b); return CAMR_ERR_PARAM; }
                    o = b;
                }
    }
    CAMR_HIP(hipMalloc(&owner_, ncell * sizeof(int)));
    CAMR_HIP(hipMemcpy(owner_, own.data(), ncell * sizeof(int), hipMemcpyHostToDevice));
    // ghost cells whose FillPatch source no box of the level owns (the
    // coarse-fine ghosts of a refined level) need exterior data
    needs_ext_ = false;
    for (int b = 0; b < nb && !needs_ext_; ++b) {
        const Box G = grow(box_[b], CAMR_NGROW);
        for (int k = G.lo[2]; k <= G.hi[2] && !needs_ext_; ++k)
            for (int j = G.lo[1]; j <= G.hi[1] && !needs_ext_; ++j)
                for (int i = G.lo[0]; i <= G.hi[0]; ++i) {
                    const int c[3] = {i, j, k};
                    int m[3];
                    for (int d = 0; d < 3; ++d) {
                        const int lo = dom.lo[d], hi = dom.hi[d], n = hi - lo + 1;
                        int x = c[d];
                        if (x < lo || x > hi) {
                            if (desc_.is_periodic[d]) x = lo + (((x - lo) % n) + n) % n;
                            else {
                                const int bc = x < lo ? desc_.hp.bclo[d] : desc_.hp.bchi[d];
                                const bool copy = bc == CAMR_BC_OUTFLOW || bc == CAMR_BC_INFLOW;
                                x = copy ? (x < lo ? lo : hi) : (x < lo ? 2 * lo - x - 1 : 2 * hi - x + 1);
                            }
                        }
                        m[d] = x;
                    }
                    if (own[(m[0] - dom.lo[0]) + (m[1] - dom.lo[1]) * n0 +
                            (int64_t)(m[2] - dom.lo[2]) * n0 * n1] < 0) {
                        needs_ext_ = true;
                        break;
                    }
                }
    }
    CAMR_HIP(hipMalloc(&tab_old_, nb * sizeof(DevFab)));
    CAMR_HIP(hipMalloc(&tab_a_, nb * sizeof(DevFab)));
    CAMR_HIP(hipMalloc(&tab_b_, nb * sizeof(DevFab)));
    {
        std::vector<DevFab> ta(nb), tb(nb);
        for (int b = 0; b < nb; ++b) {
            ta[b].a = dense(bufA_[b], box_[b]);
            ta[b].b = box_[b];
            tb[b].a = dense(bufB_[b], box_[b]);
            tb[b].b = box_[b];
        }
        CAMR_HIP(hipMemcpy(tab_a_, ta.data(), nb * sizeof(DevFab), hipMemcpyHostToDevice));
        CAMR_HIP(hipMemcpy(tab_b_, tb.data(), nb * sizeof(DevFab), hipMemcpyHostToDevice));
    }
    CAMR_HIP(hipMalloc(&tab_sb_, nb * sizeof(DevFab)));
    if (grav_src_values_) CAMR_HIP(hipMalloc(&tab_src_, nb * sizeof(DevFab)));
    std::vector<DevFab> t(nb);
    for (int b = 0; b < nb; ++b) { t[b].a = sbv_[b]; t[b].b = box_[b]; }
    CAMR_HIP(hipMemcpy(tab_sb_, t.data(), nb * sizeof(DevFab), hipMemcpyHostToDevice));
    if (grav_src_values_) {
        for (int b = 0; b < nb; ++b) t[b].a = dense(srchyd_[b], grow(box_[b], CAMR_NGROW));
        CAMR_HIP(hipMemcpy(tab_src_, t.data(), nb * sizeof(DevFab), hipMemcpyHostToDevice));
    }
    CAMR_HIP(hipMalloc(&status_, sizeof(int)));
    CAMR_HIP(hipMalloc(&redo_, sizeof(int)));
    CAMR_HIP(hipMemset(redo_, 0, sizeof(int)));
    CAMR_HIP(hipMalloc(&dtmin_, sizeof(unsigned long long)));
    CAMR_HIP(hipMalloc(&dt_dev_, sizeof(double)));
    CAMR_HIP(hipMemset(status_, 0, sizeof(int)));
    CAMR_HIP(hipHostMalloc(&h_status_, sizeof(int)));
    CAMR_HIP(hipHostMalloc(&h_dtmin_, sizeof(unsigned long long)));
    // the fused engine implements the Godunov integrator only
    fused_ = (desc_.engine != 1) && !desc_.hp.do_mol;
    batched_ = desc_.batched != 0;
    if (fused_) {
        std::vector<Arr> va(nb), vb(nb);
        for (int b = 0; b < nb; ++b) {
            va[b] = dense(bufA_[b], box_[b]);
            vb[b] = dense(bufB_[b], box_[b]);
        }
        rc = engine_.setup(box_, va, vb, default_tuning());
        // boxes beyond the fused engine's 32-bit layout run on the staged
        // engine unless the fused one was asked for explicitly
        if (rc == CAMR_ERR_SHAPE && desc_.engine == 0) fused_ = false;
        else if (rc) return rc;
    }
    if (!fused_) {
        fstride_ = max_nodes;
        {
            std::vector<DevFab> td(nb);
            for (int b = 0; b < nb; ++b) { td[b].a = dsv_[b]; td[b].b = box_[b]; }
            CAMR_HIP(hipMalloc(&tab_dsdt_, nb * sizeof(DevFab)));
            CAMR_HIP(hipMemcpy(tab_dsdt_, td.data(), nb * sizeof(DevFab), hipMemcpyHostToDevice));
        }
        for (int d = 0; d < 3; ++d)
            if ((rc = alloc(&flx_[d], NVAR * max_nodes * nworkers()))) return rc;
        size_t sc = 0;
        for (const Box &B : box_)
            sc = std::max(sc, desc_.hp.do_mol ? umdrv_mol_scratch_doubles(B)
                                              : umdrv_staged_scratch_doubles(B));
        if ((rc = setup_workers(max_g4, sc))) return rc;
    }
    return CAMR_OK;
}

// Embedded boundary (the AMREX_USE_EB build of the level): ghost width 6
// (CAMR::numGrow for Godunov + FluxRedist, CAMR.H:455-483), the cut-cell
// flags over flo..fhi (which must hold the domain grown by 7) and the box
// paths of CAMR_construct_hydro_source.cpp:80-83,150-151.  The level runs
// on the staged engine; every box needs the cut-cell data of set_box_eb when
// its kind is 1.
int HydroLevel::set_eb(const int32_t *flags, const int32_t flo[3], const int32_t fhi[3]) {
    if (!flags || !flo || !fhi) { set_error("null argument"); return CAMR_ERR_PARAM; }
    if (fused_) { set_error("embedded boundaries run on the staged engine"); return CAMR_ERR_PARAM; }
    if (desc_.hp.do_mol) {
        set_error("embedded boundaries with the method of lines (MOL_umeth_eb) are not built");
        return CAMR_ERR_PARAM;
    }
    if (ebflags_) { set_error("embedded boundary already set"); return CAMR_ERR_PARAM; }
    if (grav_src_values_) {
        // the hydro source would need the 6 ghost cells of the cut-cell path
        set_error("embedded boundaries with a constant-gravity source are not built");
        return CAMR_ERR_PARAM;
    }
    if (desc_.hp.ppm_type != 0 || desc_.hp.use_pslope != 0) {
        set_error("embedded boundaries need ppm_type = 0 and use_pslope = 0 (CAMR.cpp:186-190)");
        return CAMR_ERR_PARAM;
    }
    for (int d = 0; d < 3; ++d)
        if (std::fabs(desc_.hp.dx[d] - desc_.hp.dx[0]) > 1e-12 * desc_.hp.dx[0]) {
            set_error("embedded boundaries need dx == dy == dz (CAMR.cpp:305-313)");
            return CAMR_ERR_PARAM;
        }
    const Box fb = make_box(flo, fhi);
    const Box dom = make_box(desc_.hp.domlo, desc_.hp.domhi);
    if (!fb.ok() || !fb.contains(grow(dom, 7))) {
        set_error("EB flags must cover the domain grown by 7 cells");
        return CAMR_ERR_SHAPE;
    }
    if (needs_ext_) { set_error("embedded boundaries need boxes that cover the domain"); return CAMR_ERR_SHAPE; }
    int rc = resolve();
    if (rc) return rc;
    CAMR_HIP(hipStreamSynchronize(stream_));
    release_graphs();
    const size_t nf = fb.npts();
    CAMR_HIP(hipMalloc(&ebflags_, nf * sizeof(int32_t)));
    CAMR_HIP(hipMemcpy(ebflags_, flags, nf * sizeof(int32_t), hipMemcpyHostToDevice));
    ebf_.p = ebflags_;
    ebf_.lo0 = fb.lo[0]; ebf_.lo1 = fb.lo[1]; ebf_.lo2 = fb.lo[2];
    ebf_.js = fb.len(0);
    ebf_.ks = (int64_t)fb.len(0) * fb.len(1);
    EbFlags hf = ebf_;
    hf.p = flags;
    ng_ = 6;
    // the ghost regions grew: rebuild the exterior cell lists on the next fill
    drop_shell_cells();
    hc_.eb = 1;
    const int nb = (int)box_.size();
    kind_.assign(nb, 0);
    ebg_.assign(nb, EbGeom{});
    size_t max_g = 0, sc = 0;
    for (int b = 0; b < nb; ++b) {
        const Box B = box_[b];
        bool all_cov = true, nonreg = false;
        const Box G2 = grow(B, 2);
        for (int k = G2.lo[2]; k <= G2.hi[2]; ++k)
            for (int j = G2.lo[1]; j <= G2.hi[1]; ++j)
                for (int i = G2.lo[0]; i <= G2.hi[0]; ++i) {
                    const int t = eb_type(hf(i, j, k));
                    if (B.contains(i, j, k) && t != kEbCovered) all_cov = false;
                    if (t != kEbRegular) nonreg = true;
                }
        kind_[b] = all_cov ? 2 : (nonreg ? 1 : 0);
        ebg_[b].f = ebf_;
        max_g = std::max<size_t>(max_g, grow(B, ng_).npts());
        sc = std::max(sc, umdrv_staged_scratch_doubles(B));
        // Sborder with the wider ghost region
        (void)hipFree(sb_[b]);
        sb_[b] = nullptr;
        CAMR_HIP(hipMalloc(&sb_[b], NVAR * grow(B, ng_).npts() * sizeof(double)));
        sbv_[b] = dense(sb_[b], grow(B, ng_));
    }
    std::vector<DevFab> t(nb);
    for (int b = 0; b < nb; ++b) { t[b].a = sbv_[b]; t[b].b = box_[b]; }
    CAMR_HIP(hipMemcpy(tab_sb_, t.data(), nb * sizeof(DevFab), hipMemcpyHostToDevice));
    if ((rc = eb_regular_batch())) return rc;
    rc = setup_workers(max_g, sc);
    if (rc) return rc;
    // the cut-cell boxes: arrays per box, table entries rewritten when a
    // box's geometry arrives
    cut_ids_.clear();
    cut_off_.clear();
    cut_cells_ = nullptr;  // (an earlier list is freed with ebbuf_)
    ncut_cells_ = 0;
    cut_cells_h_.clear();
    size_t tot = 0;
    auto r32 = [](size_t n) { return (n + 31) / 32 * 32; };
    for (int b = 0; b < nb; ++b) {
        if (kind_[b] != 1) continue;
        const Box B = box_[b], G = grow(B, ng_);
        cut_ids_.push_back(b);
        cut_off_.push_back(tot);
        tot += r32((size_t)(2 * QVAR + NQAUX) * G.npts());
        for (int d = 0; d < 3; ++d) tot += r32((size_t)NVAR * nodes(B, d).npts());
        tot += r32(umdrv_eb_scratch_doubles(B, G));
    }
    if (!cut_ids_.empty()) {
        CAMR_HIP(hipMalloc(&cut_buf_, tot * sizeof(double)));
        ebbuf_.push_back(cut_buf_);
        CAMR_HIP(hipMalloc(&cut_tab_, cut_ids_.size() * sizeof(EbBox)));
        ebbuf_.push_back(reinterpret_cast<double *>(cut_tab_));
        cut_htab_.assign(cut_ids_.size(), EbBox{});
        for (size_t c = 0; c < cut_ids_.size(); ++c)
            if ((rc = cut_entry((int)c))) return rc;
        // the cut cells, for the cut-cell divergence (the cells where
        // b_eb_div<true> acts)
        std::vector<CutCell> cells;
        for (size_t c = 0; c < cut_ids_.size(); ++c) {
            const Box G = grow(box_[cut_ids_[c]], 2);
            for (int k = G.lo[2]; k <= G.hi[2]; ++k)
                for (int j = G.lo[1]; j <= G.hi[1]; ++j)
                    for (int i = G.lo[0]; i <= G.hi[0]; ++i)
                        if (eb_type(hf(i, j, k)) == kEbSingle) cells.push_back({(int)c, i, j, k});
        }
        ncut_cells_ = (int)cells.size();
        cut_cells_h_ = cells;
        if (ncut_cells_) {
            CAMR_HIP(hipMalloc(&cut_cells_, cells.size() * sizeof(CutCell)));
            ebbuf_.push_back(reinterpret_cast<double *>(cut_cells_));
            CAMR_HIP(hipMemcpy(cut_cells_, cells.data(), cells.size() * sizeof(CutCell),
                               hipMemcpyHostToDevice));
        }
    }
    wmap_.clear();
    if (eb_batched_ && nws_ > 1) {
        // the regular batch runs on worker 0: spread the cut-cell boxes over
        // the other workers
        wmap_.assign(nb, 0);
        int r = 0;
        for (int b = 0; b < nb; ++b)
            wmap_[b] = kind_[b] == 1 ? 1 + (r++ % (nws_ - 1)) : b % nws_;
    }
    return CAMR_OK;
}

// The regular boxes of an EB level (kind 0) as one batch of the fused
// kernels.  Scanning them in (z, y, x) order of their low corners, each box
// not yet taken starts the largest block of sx x sy x sz (each 1..4, most
// boxes first) untaken regular boxes of its own shape that tile a box the
// fused engine's layout holds; the batch runs that box instead of them.  A box's results depend
// only on Sborder over grow(bx, 6), whose ghost cells hold the neighbours'
// valid data (FillPatch's resync), so the virtual box computes the same
// fluxes and dS/dt at every cell.
int HydroLevel::eb_regular_batch() {
    const int nb = (int)box_.size();
    for (int pass = 0; pass < 2; ++pass) {
        const bool merge = eb_merge_ && pass == 0;
        std::vector<int> vof(nb, -1);
        std::vector<std::vector<int>> groups;  // boxes of each block, x fastest
        std::vector<Box> vb;
        std::vector<std::array<int, 3>> vsh;
        if (merge) {
            std::map<std::array<int, 3>, int> at;
            std::vector<int> ord;
            for (int b = 0; b < nb; ++b)
                if (kind_[b] == 0) {
                    at[{box_[b].lo[0], box_[b].lo[1], box_[b].lo[2]}] = b;
                    ord.push_back(b);
                }
            std::sort(ord.begin(), ord.end(), [&](int x, int y) {
                const Box &X = box_[x], &Y = box_[y];
                return std::array<int, 3>{X.lo[2], X.lo[1], X.lo[0]} <
                       std::array<int, 3>{Y.lo[2], Y.lo[1], Y.lo[0]};
            });
            std::vector<std::array<int, 3>> shapes;
            for (int sz = 1; sz <= 4; ++sz)
                for (int sy = 1; sy <= 4; ++sy)
                    for (int sx = 1; sx <= 4; ++sx)
                        if (sx * sy * sz > 1) shapes.push_back({sx, sy, sz});
            std::stable_sort(shapes.begin(), shapes.end(),
                             [](const std::array<int, 3> &x, const std::array<int, 3> &y) {
                                 return x[0] * x[1] * x[2] > y[0] * y[1] * y[2];
                             });
            for (int b : ord) {
                if (vof[b] >= 0) continue;
                const Box B = box_[b];
                const int len[3] = {B.len(0), B.len(1), B.len(2)};
                for (const auto &sh : shapes) {
                    std::vector<int> g;
                    bool ok = true;
                    for (int c = 0; c < sh[2] && ok; ++c)
                        for (int r = 0; r < sh[1] && ok; ++r)
                            for (int a = 0; a < sh[0] && ok; ++a) {
                                auto it = at.find({B.lo[0] + a * len[0], B.lo[1] + r * len[1],
                                                   B.lo[2] + c * len[2]});
                                ok = it != at.end() && vof[it->second] < 0;
                                if (!ok) break;
                                const Box M = box_[it->second];
                                for (int d = 0; d < 3; ++d) ok = ok && M.len(d) == len[d];
                                g.push_back(it->second);
                            }
                    Box V = B;
                    for (int d = 0; d < 3; ++d) V.hi[d] = B.lo[d] + sh[d] * len[d] - 1;
                    if (!ok || !FusedEngine::fits(V)) continue;
                    for (int m : g) vof[m] = (int)groups.size();
                    groups.push_back(g);
                    vb.push_back(V);
                    vsh.push_back(sh);
                    break;
                }
            }
        }
        // the virtual boxes' Sborder and dS/dt
        const int nv = (int)groups.size();
        auto r32 = [](size_t n) { return (n + 31) / 32 * 32; };
        size_t tot = 0;
        for (const Box &V : vb) tot += r32((size_t)NVAR * grow(V, ng_).npts()) + r32((size_t)NVAR * V.npts());
        double *vbuf = nullptr;
        if (nv) {
            CAMR_HIP(hipMalloc(&vbuf, tot * sizeof(double)));
            ebbuf_.push_back(vbuf);
        }
        std::vector<VGather> vg(nv);
        std::vector<Arr> vds(nv);
        std::vector<int> vids;
        double *p = vbuf;
        for (int v = 0; v < nv; ++v) {
            const Box V = vb[v];
            vg[v].v = V;
            vg[v].sb = dense(p, grow(V, ng_));
            p += r32((size_t)NVAR * grow(V, ng_).npts());
            vds[v] = dense(p, V);
            p += r32((size_t)NVAR * V.npts());
            for (int d = 0; d < 3; ++d) {
                vg[v].s[d] = vsh[v][d];
                vg[v].len[d] = box_[groups[v][0]].len(d);
            }
            vg[v].off = (int)vids.size();
            vids.insert(vids.end(), groups[v].begin(), groups[v].end());
        }
        for (int b = 0; b < nb; ++b) dsv_[b] = vof[b] >= 0 ? vds[vof[b]] : dense(dsdt_[b], box_[b]);
        // the batch: virtual boxes (rows nb + v of the extended tables), then
        // the regular boxes left alone
        std::vector<Box> rb;
        std::vector<Arr> va, vbv;
        std::vector<int> ids;
        for (int v = 0; v < nv; ++v) {
            rb.push_back(vb[v]);
            va.push_back(vds[v]);  // the state views are read by the update only, which the
            vbv.push_back(vds[v]); // batch leaves to the level
            ids.push_back(nb + v);
        }
        for (int b = 0; b < nb; ++b) {
            if (kind_[b] != 0 || vof[b] >= 0) continue;
            rb.push_back(box_[b]);
            va.push_back(dense(bufA_[b], box_[b]));
            vbv.push_back(dense(bufB_[b], box_[b]));
            ids.push_back(b);
        }
        eb_batched_ = false;
        vbox_.clear();
        vof_.assign(nb, -1);
        vgh_.clear();
        vds_.clear();
        nmerged_ = 0;
        if (rb.empty()) break;
        int rc = reg_engine_.setup(rb, va, vbv, default_tuning(), ids);
        reg_engine_.set_qgrow(ng_);
        if (rc == CAMR_ERR_SHAPE && merge) continue;  // retry box by box
        if (rc == CAMR_ERR_SHAPE) break;              // too large: per-box path
        if (rc) return rc;
        eb_batched_ = true;
        vbox_ = vb;
        vof_ = vof;
        vgh_ = vg;
        vds_ = vds;
        for (int b = 0; b < nb; ++b) nmerged_ += vof[b] >= 0;
        // the boxes of a virtual box read and write its Sborder
        for (int b = 0; b < nb; ++b)
            if (vof[b] >= 0) sbv_[b] = vg[vof[b]].sb;
        std::vector<DevFab> ts(nb + nv), td(nb + nv);
        for (int b = 0; b < nb; ++b) {
            ts[b].a = sbv_[b]; ts[b].b = box_[b];
            td[b].a = dsv_[b]; td[b].b = box_[b];
        }
        for (int v = 0; v < nv; ++v) {
            ts[nb + v].a = vg[v].sb; ts[nb + v].b = vb[v];
            td[nb + v].a = vds[v]; td[nb + v].b = vb[v];
        }
        CAMR_HIP(hipMalloc(&tab_sbx_, (nb + nv) * sizeof(DevFab)));
        ebbuf_.push_back(reinterpret_cast<double *>(tab_sbx_));
        CAMR_HIP(hipMalloc(&tab_dsdtx_, (nb + nv) * sizeof(DevFab)));
        ebbuf_.push_back(reinterpret_cast<double *>(tab_dsdtx_));
        CAMR_HIP(hipMemcpy(tab_sbx_, ts.data(), ts.size() * sizeof(DevFab), hipMemcpyHostToDevice));
        CAMR_HIP(hipMemcpy(tab_dsdtx_, td.data(), td.size() * sizeof(DevFab), hipMemcpyHostToDevice));
        CAMR_HIP(hipMemcpy(tab_dsdt_, td.data(), nb * sizeof(DevFab), hipMemcpyHostToDevice));
        CAMR_HIP(hipMemcpy(tab_sb_, ts.data(), nb * sizeof(DevFab), hipMemcpyHostToDevice));
        if (nv) {
            // expand_state's rows: the boxes left alone, then the virtual boxes
            std::vector<DevFab> te;
            for (int b = 0; b < nb; ++b)
                if (vof[b] < 0) { te.push_back(ts[b]); exp_box_.push_back(box_[b]); }
            for (int v = 0; v < nv; ++v) { te.push_back(ts[nb + v]); exp_box_.push_back(vb[v]); }
            CAMR_HIP(hipMalloc(&tab_exp_, te.size() * sizeof(DevFab)));
            ebbuf_.push_back(reinterpret_cast<double *>(tab_exp_));
            CAMR_HIP(hipMemcpy(tab_exp_, te.data(), te.size() * sizeof(DevFab), hipMemcpyHostToDevice));
            CAMR_HIP(hipMalloc(&vg_ids_, vids.size() * sizeof(int)));
            ebbuf_.push_back(reinterpret_cast<double *>(vg_ids_));
            CAMR_HIP(hipMemcpy(vg_ids_, vids.data(), vids.size() * sizeof(int), hipMemcpyHostToDevice));
        }
        break;
    }
    if (!eb_batched_) {
        for (int b = 0; b < nb; ++b) dsv_[b] = dense(dsdt_[b], box_[b]);
        std::vector<DevFab> td(nb);
        for (int b = 0; b < nb; ++b) { td[b].a = dsv_[b]; td[b].b = box_[b]; }
        CAMR_HIP(hipMemcpy(tab_dsdt_, td.data(), nb * sizeof(DevFab), hipMemcpyHostToDevice));
    }
    return CAMR_OK;
}

// The force on the embedded body, F = sum over the cut cells of the wall
// momentum flux (compute_hyp_wallflux's fluxw[UMX..UMZ],
// Hydro_hypflux_eb_K.H:65) times dx^2, from the current S_new.  CAMR itself
// computes no force; this is the functional of the Sod-past-a-sphere
// workload.  Summed on the host box by box, cell by cell (x fastest).
int HydroLevel::eb_force(double f[3]) {
    if (!ebflags_) { set_error("the level has no embedded boundary"); return CAMR_ERR_PARAM; }
    int rc = resolve();
    if (rc) return rc;
    f[0] = f[1] = f[2] = 0.0;
    const double dx2 = desc_.hp.dx[0] * desc_.hp.dx[0];
    std::vector<double> h;
    for (size_t b = 0; b < box_.size(); ++b) {
        if (kind_[b] != 1) continue;
        if (!ebg_[b].vf.p) { set_error("box %d has cut cells but no cut-cell data", (int)b); return CAMR_ERR_PARAM; }
        const Box B = box_[b];
        const size_t np = B.npts();
        if (sstride_ < 3 * np) { set_error("scratch too small for the force"); return CAMR_ERR_PARAM; }
        Arr q = dense(q_, B), qa = dense(qaux_, B), out = dense(scratch_.base(), B);
        launch_on(k_ctoprim, B, stream_, snew_[b], q, qa, hc_, status_);
        if ((rc = eb_force_box(B, q, qa, ebg_[b], hc_, out, dx2, stream_))) return rc;
        h.resize(3 * np);
        CAMR_HIP(hipStreamSynchronize(stream_));
        CAMR_HIP(hipMemcpy(h.data(), scratch_.base(), 3 * np * sizeof(double), hipMemcpyDeviceToHost));
        for (int d = 0; d < 3; ++d)
            for (size_t c = 0; c < np; ++c) f[d] += h[(size_t)d * np + c];
    }
    return check_status();
}

// q / qaux / srcq (max_g points each) and scratch (sc doubles) per worker
// stream; with a single box the level's own stream does all the work.
int HydroLevel::setup_workers(size_t max_g, size_t sc) {
    auto fr = [](double *&p) { if (p) (void)hipFree(p); p = nullptr; };
    fr(q_); fr(qaux_); fr(srcq_);
    const int nw = (box_.size() > 1 && !grav_src_values_) ? nworkers() : 1;
    qstride_ = max_g;
    sstride_ = (sc + 63) / 64 * 64;
    CAMR_HIP(hipMalloc(&q_, QVAR * max_g * nw * sizeof(double)));
    CAMR_HIP(hipMalloc(&qaux_, NQAUX * max_g * nw * sizeof(double)));
    CAMR_HIP(hipMalloc(&srcq_, QVAR * max_g * nw * sizeof(double)));
    if (!scratch_.reserve(sstride_ * nw * sizeof(double))) {
        set_error("scratch allocation of %zu bytes failed", sstride_ * nw * sizeof(double));
        return CAMR_ERR_ALLOC;
    }
    if (nw > 1 && !nws_) {
        for (int w = 0; w < nw; ++w) {
            CAMR_HIP(hipStreamCreateWithFlags(&wst_[w], hipStreamNonBlocking));
            CAMR_HIP(hipEventCreateWithFlags(&wev_[w], hipEventDisableTiming));
        }
        CAMR_HIP(hipEventCreateWithFlags(&mev_, hipEventDisableTiming));
        nws_ = nw;
    }
    return CAMR_OK;
}

// Worker streams wait for the level's stream / the level's stream waits for
// every worker.
int HydroLevel::fork() {
    if (!nws_) return CAMR_OK;
    CAMR_HIP(hipEventRecord(mev_, stream_));
    for (int w = 0; w < nws_; ++w) CAMR_HIP(hipStreamWaitEvent(wst_[w], mev_, 0));
    return CAMR_OK;
}
int HydroLevel::join() {
    if (!nws_) return CAMR_OK;
    for (int w = 0; w < nws_; ++w) {
        CAMR_HIP(hipEventRecord(wev_[w], wst_[w]));
        CAMR_HIP(hipStreamWaitEvent(stream_, wev_[w], 0));
    }
    return CAMR_OK;
}

int HydroLevel::set_box_eb(int b, const double *vf, const double *const ap[3],
                           const double *const fc[3]) {
    if (!ebflags_) { set_error("set the embedded boundary first"); return CAMR_ERR_PARAM; }
    if (b < 0 || b >= (int)box_.size()) { set_error("bad box index"); return CAMR_ERR_PARAM; }
    if (!vf || !ap[0] || !ap[1] || !ap[2] || !fc[0] || !fc[1] || !fc[2]) {
        set_error("null argument");
        return CAMR_ERR_PARAM;
    }
    const Box G = grow(box_[b], 7);
    release_graphs();  // captured steps hold the geometry pointers
    auto up = [&](const double *h, const Box &bb, int nc, Arr *out) -> int {
        double *p = nullptr;
        const size_t n = (size_t)nc * bb.npts();
        CAMR_HIP(hipMalloc(&p, n * sizeof(double)));
        ebbuf_.push_back(p);
        CAMR_HIP(hipMemcpy(p, h, n * sizeof(double), hipMemcpyHostToDevice));
        *out = dense(p, bb);
        return CAMR_OK;
    };
    EbGeom &g = ebg_[b];
    int rc = up(vf, G, 1, &g.vf);
    for (int d = 0; d < 3 && !rc; ++d) rc = up(ap[d], nodes(G, d), 1, &g.ap[d]);
    for (int d = 0; d < 3 && !rc; ++d) rc = up(fc[d], nodes(G, d), 2, &g.fc[d]);
    g.f = ebf_;
    for (size_t c = 0; c < cut_ids_.size() && !rc; ++c)
        if (cut_ids_[c] == b) rc = cut_entry((int)c);
    return rc;
}

// Table entry c of the cut-cell batch: box cut_ids_[c] with its arrays in
// cut_buf_ (q, qaux, srcq over grow(bx, 6), the face fluxes, the scratch),
// its Sborder, dS/dt and geometry; copied to the device table.
int HydroLevel::cut_entry(int c) {
    const int b = cut_ids_[c];
    const Box B = box_[b], G = grow(B, ng_);
    auto r32 = [](size_t n) { return (n + 31) / 32 * 32; };
    double *p = cut_buf_ + cut_off_[c];
    UmdrvEbArgs A;
    A.bx = B;
    A.qbox = G;
    A.hc = hc_;
    for (int d = 0; d < 3; ++d) A.periodic[d] = desc_.is_periodic[d];
    A.dtp = dt_dev_;
    A.uin = sbv_[b];
    A.dsdt = dsv_[b];
    A.q = dense(p, G);
    A.srcq = dense(p + (size_t)QVAR * G.npts(), G);
    A.qaux = dense(p + (size_t)2 * QVAR * G.npts(), G);
    p += r32((size_t)(2 * QVAR + NQAUX) * G.npts());
    for (int d = 0; d < 3; ++d) {
        A.flx[d] = dense(p, nodes(B, d));
        p += r32((size_t)NVAR * nodes(B, d).npts());
    }
    A.has_src = 0;
    A.g = ebg_[b];
    eb_box_setup(A, p, &cut_htab_[c]);
    CAMR_HIP(hipMemcpy(cut_tab_ + c, &cut_htab_[c], sizeof(EbBox), hipMemcpyHostToDevice));
    return CAMR_OK;
}

int HydroLevel::box_kind(int b, int32_t *kind) const {
    if (b < 0 || b >= (int)box_.size() || !kind) { set_error("bad argument"); return CAMR_ERR_PARAM; }
    *kind = kind_.empty() ? 0 : kind_[b];
    return CAMR_OK;
}

int HydroLevel::upload(int b, const double *h) {
    if (b < 0 || b >= (int)box_.size()) { set_error("bad box index"); return CAMR_ERR_PARAM; }
    int rc = resolve();
    if (rc) return rc;
    dt_ready_ = false;
    CAMR_HIP(hipMemcpy(snew_[b].p, h, NVAR * box_[b].npts() * sizeof(double), hipMemcpyHostToDevice));
    return CAMR_OK;
}

int HydroLevel::download(int b, double *h) {
    if (b < 0 || b >= (int)box_.size()) { set_error("bad box index"); return CAMR_ERR_PARAM; }
    int rc = resolve();
    if (rc) return rc;
    CAMR_HIP(hipStreamSynchronize(stream_));
    CAMR_HIP(hipMemcpy(h, snew_[b].p, NVAR * box_[b].npts() * sizeof(double), hipMemcpyDeviceToHost));
    return CAMR_OK;
}

int HydroLevel::state_array(int b, camr_array4 *out) {
    if (b < 0 || b >= (int)box_.size()) { set_error("bad box index"); return CAMR_ERR_PARAM; }
    const Box B = box_[b];
    out->p = snew_[b].p;
    out->jstride = B.len(0);
    out->kstride = (int64_t)B.len(0) * B.len(1);
    out->nstride = B.npts();
    for (int d = 0; d < 3; ++d) { out->lo[d] = B.lo[d]; out->hi[d] = B.hi[d]; }
    out->ncomp = NVAR;
    return CAMR_OK;
}

int HydroLevel::check_status() {
    CAMR_HIP(hipMemcpyAsync(h_status_, status_, sizeof(int), hipMemcpyDeviceToHost, stream_));
    CAMR_HIP(hipStreamSynchronize(stream_));
    return status_error(*h_status_);
}

// The error of the device status word `st` (already read back): the first
// error bit, which is then cleared on the device; the floor bit is not an
// error.
int HydroLevel::status_error(int st) {
    const int s = st & ~kFloorFlag;
    if (s) {
        CAMR_HIP(hipMemsetAsync(status_, 0, sizeof(int), stream_));
        for (int e = 1; e < 31; ++e)
            if (s & (1 << e)) {
                static const char *msg[] = {"", "", "",
                    "rho E or rho e not positive (Hydro_ctoprim.H:56,68)",
                    "density exactly zero in enforce_minimum_density (CAMR.cpp:990-992)",
                    "dt needs to be positive (Timestep.H:59)", "", "",
                    "adjoint: a state took a branch the adjoint does not differentiate "
                    "(small-energy reset)"};
                set_error("%s", (e < 9 && msg[e][0]) ? msg[e] : "device error");
                return e;
            }
    }
    return CAMR_OK;
}

int HydroLevel::sync() {
    int rc = resolve();
    if (rc) return rc;
    CAMR_HIP(hipStreamSynchronize(stream_));
    return check_status();
}

int HydroLevel::init_post() {
    dt_ready_ = false;
    for (size_t b = 0; b < box_.size(); ++b)
        launch_on(k_init_post, box_[b], stream_, snew_[b], hc_);
    CAMR_HIP(hipGetLastError());
    return sync();
}

// clean_state on valid cells (CAMR.cpp:1253-1262 with ng = 0):
// enforce_min_density -> normalize_species -> computeTemp.
int HydroLevel::clean_state(std::vector<Arr> &S, bool launch_detect) {
    for (size_t b = 0; b < box_.size(); ++b) {
        const hipStream_t s = cur_ws(b);
        if (launch_detect) launch_on(k_detect, box_[b], s, S[b], hc_.small_dens, status_, ebf_);
        hipLaunchKernelGGL(k_floor_seq, dim3(1), dim3(1), 0, s, box_[b], S[b], hc_, status_, ebf_);
        launch_on(k_clean_point, box_[b], s, S[b], hc_);
    }
    return CAMR_OK;
}

// expand_state (CAMR.cpp:1243-1250): FillPatch from S_old, then
// clean_state(Sborder) with ng = numGrow.
int HydroLevel::expand_state() {
    std::vector<DevFab> t(box_.size());
    for (size_t b = 0; b < box_.size(); ++b) { t[b].a = sold_[b]; t[b].b = box_[b]; }
    // every box's S_old is final (the caller joined the workers)
    CAMR_HIP(hipMemcpyAsync(tab_old_, t.data(), t.size() * sizeof(DevFab), hipMemcpyHostToDevice, stream_));
    if (in_ws_) fork();
    for (size_t b = 0; b < box_.size(); ++b) {
        const Box G = grow(box_[b], ng_);
        launch_on(k_fill_raw, G, cur_ws(b), sbv_[b], (const DevFab *)tab_old_, (const int *)owner_, dom_,
                  extv_, ext_ ? 1 : 0, status_);
    }
    for (size_t b = 0; b < box_.size(); ++b) {
        const hipStream_t s = cur_ws(b);
        launch_on(k_detect, box_[b], s, sbv_[b], hc_.small_dens, status_, ebf_);
        hipLaunchKernelGGL(k_floor_seq, dim3(1), dim3(1), 0, s, box_[b], sbv_[b], hc_, status_, ebf_);
        launch_on(k_normalize, box_[b], s, sbv_[b]);
    }
    // FillBoundary reads the other boxes' cleaned valid cells
    if (in_ws_) { join(); fork(); }
    for (size_t b = 0; b < box_.size(); ++b) {
        const Box G = grow(box_[b], ng_);
        launch_on(k_resync, G, cur_ws(b), box_[b], sbv_[b], (const DevFab *)tab_sb_, (const int *)owner_, dom_);
    }
    for (size_t b = 0; b < box_.size(); ++b)
        launch_on(k_resetT, grow(box_[b], ng_), cur_ws(b), sbv_[b], hc_);
    // hipMemcpyAsync from pageable host memory is synchronous w.r.t. the
    // host buffer, so `t` may go out of scope here.
    return CAMR_OK;
}

// construct_hydro_source (Hydro/CAMR_construct_hydro_source.cpp:9-267):
// per box ctoprim, srctoprim, hydro_umdrv, then S_new += dt * dSdt.
int HydroLevel::construct_hydro_source(double dt, bool add) {
    if (grav_src_values_) {
        for (size_t b = 0; b < box_.size(); ++b)
            launch_on(k_srchyd, grow(box_[b], CAMR_NGROW), stream_, box_[b],
                      dense(srchyd_[b], grow(box_[b], CAMR_NGROW)), sold_[b], desc_.const_grav);
        for (size_t b = 0; b < box_.size(); ++b)
            launch_on(k_resync, grow(box_[b], CAMR_NGROW), stream_, box_[b],
                      dense(srchyd_[b], grow(box_[b], CAMR_NGROW)), (const DevFab *)tab_src_,
                      (const int *)owner_, dom_);
    }
    if (fused_) {
        const int has_src = grav_src_values_ ? 1 : 0;
        const DevFab *src = grav_src_values_ ? (const DevFab *)tab_src_ : nullptr;
        if (batched_)
            return engine_.hydro(-1, parity(), false, tab_sb_, src, has_src, hc_, dt_dev_,
                                 stream_, status_, &prof_);
        for (size_t b = 0; b < box_.size(); ++b) {
            int rc = engine_.hydro((int)b, parity(), false, tab_sb_, src, has_src, hc_, dt_dev_,
                                   stream_, status_, &prof_);
            if (rc) return rc;
        }
        return CAMR_OK;
    } else {
        if (eb_batched_) {
            // the regular boxes of an EB level in batched launches, on the
            // first worker stream like the per-box work, so that the join
            // orders them before the update
            const hipStream_t es = in_ws_ && nws_ ? wst_[0] : stream_;
            const int has_src = grav_src_values_ ? 1 : 0;
            const DevFab *src = grav_src_values_ ? (const DevFab *)tab_src_ : nullptr;
            int rc = reg_engine_.hydro(-1, parity(), false, tab_sbx_, src, has_src, hc_, dt_dev_,
                                       es, status_, &prof_, nullptr, true);
            if (!rc) rc = reg_engine_.dsdt(-1, parity(), tab_dsdtx_, hc_, es, &prof_);
            if (rc) { set_error("kernel launch failed"); return rc; }
        }
        if (!cut_ids_.empty()) {
            // the cut-cell boxes: primitives per box, then hydro_umdrv_eb
            // (:182-226) in batched launches over all of them, on a worker
            // stream of their own
            // (an EB level has no gravity source: set_eb refuses it)
            const hipStream_t cs = in_ws_ && nws_ ? wst_[nws_ > 1 ? 1 : 0] : stream_;
            for (size_t c = 0; c < cut_ids_.size(); ++c) {
                const int b = cut_ids_[c];
                if (!ebg_[b].vf.p) {
                    set_error("box %d has cut cells but no cut-cell data (camr_level_set_box_eb)", b);
                    return CAMR_ERR_PARAM;
                }
            }
            const int nc = (int)cut_ids_.size();
            int rc = eb_prim_batch(cut_tab_, cut_htab_.data(), nc, hc_, status_, cs, &prof_);
            if (!rc)
                rc = umdrv_eb_batch(cut_tab_, cut_htab_.data(), nc, hc_, dt_dev_, 0, cs, &prof_,
                                    cut_cells_, ncut_cells_);
            if (rc) { set_error("kernel launch failed"); return rc; }
        }
        for (size_t b = 0; b < box_.size(); ++b) {
            const Box B = box_[b], G = grow(B, ng_);
            const hipStream_t bs = cur_ws(b);
            const int w = in_ws_ ? wslot(b) : 0;
            double *wq = q_ + (size_t)w * QVAR * qstride_, *wqa = qaux_ + (size_t)w * NQAUX * qstride_;
            double *wsq = srcq_ + (size_t)w * QVAR * qstride_;
            double *wscr = scratch_.base() + (size_t)w * sstride_;
            Arr ds = dsv_[b];
            const int kind = kind_.empty() ? 0 : kind_[b];
            // regular box of an EB level: reg_engine_ has written its dS/dt;
            // cut-cell box: the batch above has
            if ((kind == 0 && eb_batched_) || kind == 1) continue;
            if (kind == 2) {  // covered box: no hydro (CAMR_construct_hydro_source.cpp:80)
                launch_on(k_fill_value, B, bs, ds, (int)NVAR, 0.0);
                continue;
            }
            Arr q = dense(wq, G), qa = dense(wqa, G), sq = dense(wsq, G);
            launch_on(k_ctoprim, G, bs, sbv_[b], q, qa, hc_, status_);
            // primitive sources feed the Godunov predictor only (:129-143)
            if (grav_src_values_ && !desc_.hp.do_mol)
                launch_on(k_srctoprim, G, bs, q, qa, dense(srchyd_[b], G), sq);
            UmdrvArgs A;
            A.bx = B;
            A.qbox = G;
            A.hc = hc_;
            A.dtp = dt_dev_;
            A.uin = sbv_[b];
            A.dsdt = ds;
            A.q = q;
            A.qaux = qa;
            A.srcq = sq;
            A.has_src = grav_src_values_ ? 1 : 0;
            for (int d = 0; d < 3; ++d) {
                A.flx[d] = dense(flx_[d] + (size_t)w * NVAR * fstride_, nodes(B, d));
                A.area[d] = A.flx[d];
            }
            A.has_area = 0;
            A.area_c[0] = hc_.dx[1] * hc_.dx[2];
            A.area_c[1] = hc_.dx[0] * hc_.dx[2];
            A.area_c[2] = hc_.dx[0] * hc_.dx[1];
            A.vol = ds;
            A.has_vol = 0;
            A.vol_c = hc_.dx[0] * hc_.dx[1] * hc_.dx[2];
            // the hyd_src setVal(0) (:59) and hydro_update's += onto it, as
            // the update writing 0 + the divergence into every valid cell
            A.dsdt_zero = 1;
            int rc = desc_.hp.do_mol ? umdrv_mol(A, wscr, bs, &prof_) : umdrv_staged(A, wscr, bs, &prof_);
            if (rc) { set_error("kernel launch failed"); return rc; }
            // do_reflux: the box's fluxes into the flux registers (:243-254)
            if (reg_crse_ || reg_fine_) reg_add(b, A.flx, bs);
        }
    }
    if (add)
        for (size_t b = 0; b < box_.size(); ++b)
            launch_on(k_add_hydro, box_[b], cur_ws(b), snew_[b], dsv_[b], dt,
                      hc_.small_dens, status_, ebf_);
    return CAMR_OK;
}

int HydroLevel::advance(double dt) {
    if (!(dt > 0.0)) { set_error("dt must be positive"); return CAMR_ERR_DT; }
    if (needs_ext_ && !ext_) {
        set_error("ghost cells outside every box of the level: set exterior data");
        return CAMR_ERR_SHAPE;
    }
    int rc = resolve();
    if (rc) return rc;
    Region step(&prof_, "advance", stream_);
    // swapTimeLevels (CAMR_advance.cpp:72-75)
    std::swap(snew_, sold_);
    a_is_new_ = !a_is_new_;
    set_dt_dev(dt);
    if (fused_ && !grav_src_values_) {
        rc = advance_fused(dt);
        pending_ = true;
        last_dt_ = dt;
    } else if (use_staged_graph() && !prof_.enabled()) {
        rc = replay_staged(dt);
        dt_ready_ = staged_dt();
    } else {
        rc = advance_exact(dt);
        dt_ready_ = !fused_ && staged_dt();
    }
    if (rc) return rc;
    CAMR_HIP(hipGetLastError());
    time_ += dt;
    return CAMR_OK;
}

// The reference sequence pass by pass (CAMR_advance.cpp:72-164), starting
// after the swap; S_old holds the raw previous state.
int HydroLevel::advance_exact(double dt) {
    if (!fused_) return advance_staged(dt);
    // clean_state(S_old) (:83)
    clean_state(sold_, true);
    // expand_state(Sborder, prev_time, numGrow()) (:99)
    int rc = expand_state();
    if (rc) return rc;
    // MultiFab::Copy(S_new, Sborder) + old-source Saxpy (:112-118)
    for (size_t b = 0; b < box_.size(); ++b)
        launch_on(k_copy_oldsrc, box_[b], cur_ws(b), snew_[b], sbv_[b], sold_[b], dt,
                  desc_.const_grav, desc_.add_grav_src);
    // construct_hydro_source + Saxpy (:138-141)
    scratch_.reset();
    rc = construct_hydro_source(dt);
    if (rc) return rc;
    // clean_state(S_new) (:146)
    clean_state(snew_, false);
    if (desc_.add_grav_src) {
        // new-time source correction + clean_state (:150-154)
        for (size_t b = 0; b < box_.size(); ++b)
            launch_on(k_add_newsrc, box_[b], cur_ws(b), snew_[b], sold_[b], dt, desc_.const_grav,
                      hc_.small_dens, status_, ebf_);
        clean_state(snew_, false);
    }
    // post_timestep: clean_state(S_new) (CAMR.cpp:698-699); a coarse level
    // of a subcycled step runs it after reflux and avgDown
    if (!defer_post_) clean_state(snew_, true);
    return CAMR_OK;
}

template <typename K, typename... A>
void HydroLevel::launch_all(K kern, int g, A... args) {
    int64_t n = 1;
    for (const Box &B : box_) n = std::max<int64_t>(n, grow(B, g).npts());
    hipLaunchKernelGGL(kern, dim3(grid_for(n), (unsigned)box_.size()), dim3(kBlock), 0, stream_,
                       args...);
}

// clean_state on the valid cells of every box of `tab` (CAMR.cpp:1253-1262,
// ng = 0): three launches for the whole level.
int HydroLevel::clean_all(const DevFab *tab, bool launch_detect) {
    if (launch_detect) launch_all(kb_detect, 0, tab, hc_.small_dens, status_, ebf_);
    hipLaunchKernelGGL(kb_floor_seq, dim3(1), dim3(1), 0, stream_, tab, (int)box_.size(), hc_,
                       status_, ebf_);
    launch_all(kb_clean_point, 0, tab, hc_);
    return CAMR_OK;
}

// The staged engine's step, pass by pass as in CAMR_advance.cpp:72-164:
// the pointwise passes run as one launch over all boxes on the level's
// stream, the per-box hydro (hydro_umdrv / hydro_umdrv_eb) on the worker
// streams between a fork and a join.
int HydroLevel::advance_staged(double dt) {
    const DevFab *tnew = a_is_new_ ? tab_a_ : tab_b_, *told = a_is_new_ ? tab_b_ : tab_a_;
    // clean_state(S_old) (:83)
    clean_all(told, true);
    // expand_state(Sborder, prev_time) (:99)
    expand_all(told, true);
    scratch_.reset();
    if (desc_.hp.do_mol) {
        // first stage: the hydro source (it reads Sborder, not S_new), then
        // MultiFab::Copy(S_new, Sborder) + old-source Saxpy (:112-118) and
        // S_new += dt dS/dt (:128) in one pass
        int rc = hydro_on_workers(dt);
        if (rc) return rc;
        launch_all(kb_copy_add_hydro, 0, tnew, (const DevFab *)tab_sb_, told,
                   (const DevFab *)tab_dsdt_, (const double *)dt_dev_, desc_.const_grav,
                   desc_.add_grav_src, hc_.small_dens, status_, ebf_);
        // second stage (:130-134): expand_state(Sborder, cur_time) from
        // S_new and the hydro source of that state; the combination
        // S_new = 0.5 Sborder + 0.5 S_old + 0.5 dt dSdt opens the tail
        expand_all(tnew);
        rc = hydro_on_workers(dt);
        if (rc) return rc;
        staged_tail(tnew, told, 1);
        return CAMR_OK;
    }
    // construct_hydro_source (:138) per box on the worker streams (it reads
    // Sborder, not S_new)
    int rc = hydro_on_workers(dt);
    if (rc) return rc;
    staged_tail(tnew, told, 0);
    return CAMR_OK;
}

// The step's pointwise tail after the hydro source in one pass: S_new
// (Sborder + old source + dt dS/dt, :112-141; with mol the second MOL
// stage, :134), clean_state (:146), the new-time source and its
// clean_state (:150-154) and post_timestep's clean_state (CAMR.cpp:
// 698-699), then the CFL estimate of the result; repeated pass by pass by
// kb_tail_redo where a density floor would act.
void HydroLevel::staged_tail(const DevFab *tnew, const DevFab *told, int mol) {
    const int post = defer_post_ ? 0 : 1;
    // with post_timestep's clean the state is final: its CFL estimate
    // (estTimeStep's minimum) comes out of the same pass
    unsigned long long *dtm = post ? dtmin_ : nullptr;
    if (dtm)
        launch_on(k_fill_u64, make_box1(), stream_, dtmin_,
                  (unsigned long long)0x7fefffffffffffffULL);
    launch_all(kb_tail_fused, 0, tnew, (const DevFab *)tab_sb_, told, (const DevFab *)tab_dsdt_,
               (const double *)dt_dev_, desc_.const_grav, desc_.add_grav_src, post, mol, hc_,
               redo_, ebf_, dtm, status_);
    hipLaunchKernelGGL(kb_tail_redo, dim3(1), dim3(1024), 0, stream_, tnew,
                       (const DevFab *)tab_sb_, told, (const DevFab *)tab_dsdt_,
                       (int)box_.size(), (const double *)dt_dev_, desc_.const_grav,
                       desc_.add_grav_src, post, mol, hc_, status_, redo_, ebf_, dtm);
}

// expand_state (CAMR.cpp:1243-1250) from the states of `src`: FillPatch,
// clean_state(Sborder) on the valid cells, FillBoundary, computeTemp on
// the grown boxes.
void HydroLevel::expand_all(const DevFab *src, bool cleaned) {
    if (cleaned && tab_exp_) {
        // virtual boxes as one row each, the rows' blocks in one dimension
        FlatBlocks fb;
        const dim3 grid = flat_blocks(
            (int)exp_box_.size(),
            [&](int r) -> int64_t { return grid_for(grow(exp_box_[r], ng_).npts()); }, fb);
        hipLaunchKernelGGL(kb_expand_cleaned, grid, dim3(kBlock), 0, stream_,
                           (const DevFab *)tab_exp_, ng_, src, (const int *)owner_, dom_, extv_,
                           ext_ ? 1 : 0, hc_, status_, ebf_, fb, (int *)nullptr);
        return;
    }
    if (!tab_exp_) {
        // one pass; from an uncleaned src, repeated pass by pass where the
        // floor pass between the detection and normalize_species would act
        FlatBlocks fb;
        fb.n = 0;  // boxes of the level: blockIdx.y
        int *rd = cleaned ? nullptr : redo_;
        launch_all(kb_expand_cleaned, ng_, (const DevFab *)tab_sb_, ng_, src, (const int *)owner_,
                   dom_, extv_, ext_ ? 1 : 0, hc_, status_, ebf_, fb, rd);
        if (rd)
            hipLaunchKernelGGL(kb_expand_redo, dim3(1), dim3(1024), 0, stream_,
                               (const DevFab *)tab_sb_, (int)box_.size(), ng_, src,
                               (const int *)owner_, dom_, extv_, ext_ ? 1 : 0, hc_, status_, ebf_,
                               rd);
        return;
    }
    launch_all(kb_fill_raw, ng_, (const DevFab *)tab_sb_, ng_, src, (const int *)owner_, dom_,
               extv_, ext_ ? 1 : 0, status_);
    launch_all(kb_detect, 0, (const DevFab *)tab_sb_, hc_.small_dens, status_, ebf_);
    hipLaunchKernelGGL(kb_floor_seq, dim3(1), dim3(1), 0, stream_, (const DevFab *)tab_sb_,
                       (int)box_.size(), hc_, status_, ebf_);
    launch_all(kb_normalize, 0, (const DevFab *)tab_sb_);
    launch_all(kb_resync, ng_, (const DevFab *)tab_sb_, ng_, (const DevFab *)tab_sb_,
               (const int *)owner_, dom_);
    launch_all(kb_resetT, ng_, (const DevFab *)tab_sb_, ng_, hc_);
}

// construct_hydro_source into dSdt of every box, spread over the worker
// streams between a fork and a join.
int HydroLevel::hydro_on_workers(double dt) {
    int rc = fork();
    if (rc) return rc;
    in_ws_ = true;
    rc = construct_hydro_source(dt, false);
    in_ws_ = false;
    const int rj = join();
    return rc ? rc : rj;
}

// Same step with the pointwise passes fused (level_fused.hip); exact unless a
// density floor fires, which resolve() detects and answers with
// advance_exact from the untouched raw S_old.
int HydroLevel::advance_fused(double dt) {
    (void)dt;
    const unsigned long long init = (unsigned long long)0x7fefffffffffffffULL;  // DBL_MAX
    launch_on(k_fill_u64, make_box1(), stream_, dtmin_, init);
    FillArgs fa;
    fa.sold_tab = a_is_new_ ? tab_b_ : tab_a_;  // table of the buffers now old
    fa.owner = owner_;
    fa.ext = extv_;
    fa.has_ext = ext_ ? 1 : 0;
    fa.dom = dom_;
    fa.hc = hc_;
    fa.dtp = dt_dev_;
    fa.g = desc_.const_grav;
    fa.add_grav = desc_.add_grav_src;
    PostArgs pa;
    pa.hc = hc_;
    pa.dtp = dt_dev_;
    pa.g = desc_.const_grav;
    pa.add_grav = desc_.add_grav_src;
    pa.defer_post = defer_post_ ? 1 : 0;
    pa.dtmin = dtmin_;
    const int nb = batched_ ? 1 : (int)box_.size();
    for (int i = 0; i < nb; ++i) {
        const int b = batched_ ? -1 : i;
        int rc = engine_.fill(b, parity(), fa, stream_, status_, &prof_);
        if (!rc) rc = engine_.hydro(b, parity(), true, nullptr, nullptr, 0, hc_, dt_dev_, stream_,
                                    status_, &prof_, &pa);
        if (rc) return rc;
    }
    dt_ready_ = true;
    return CAMR_OK;
}

void HydroLevel::release_graphs() {
    for (int p = 0; p < 2; ++p) {
        if (gexec_[p]) (void)hipGraphExecDestroy(gexec_[p]);
        if (graph_[p]) (void)hipGraphDestroy(graph_[p]);
        gexec_[p] = nullptr;
        graph_[p] = nullptr;
        for (int q = 0; q < 2; ++q) {
            if (sgexec_[p][q]) (void)hipGraphExecDestroy(sgexec_[p][q]);
            if (sgraph_[p][q]) (void)hipGraphDestroy(sgraph_[p][q]);
            sgexec_[p][q] = nullptr;
            sgraph_[p][q] = nullptr;
        }
    }
}

// One graph launch per step on the staged engine: the whole pass sequence
// of advance_staged (every launch reads dt from dt_dev_, the tables and
// scratch are fixed), captured on first use for the current buffer parity.
int HydroLevel::replay_staged(double dt) {
    const int p = parity();
    const int q = defer_post_ ? 1 : 0;
    if (!sgexec_[p][q]) {
        // the legacy default stream cannot be captured: record on the
        // level's capture stream, replay on its own stream
        if (!gstream_) CAMR_HIP(hipStreamCreateWithFlags(&gstream_, hipStreamNonBlocking));
        CAMR_HIP(hipStreamSynchronize(stream_));
        const hipStream_t own = stream_;
        stream_ = gstream_;
        hipError_t e = hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal);
        int rc = e == hipSuccess ? advance_staged(dt) : CAMR_ERR_HIP;
        hipGraph_t g = nullptr;
        if (e == hipSuccess) e = hipStreamEndCapture(stream_, &g);
        stream_ = own;
        if (rc || e != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            if (!rc) set_error("capture of the level step failed: %s", hipGetErrorString(e));
            return rc ? rc : CAMR_ERR_HIP;
        }
        sgraph_[p][q] = g;
        CAMR_HIP(hipGraphInstantiate(&sgexec_[p][q], g, nullptr, nullptr, 0));
    }
    CAMR_HIP(hipGraphLaunch(sgexec_[p][q], stream_));
    return CAMR_OK;
}

// Capture one level step per buffer parity: device dt, fill, the hydro
// stages and the post pass, every stage one launch over all boxes, reading
// dt from the control block.  No memset or host work inside.
int HydroLevel::capture_graphs() {
    if (!gstream_) CAMR_HIP(hipStreamCreateWithFlags(&gstream_, hipStreamNonBlocking));
    DtParams P;
    P.fixed_dt = desc_.fixed_dt;
    P.max_dt = desc_.max_dt;
    P.cfl = desc_.cfl;
    P.change_max = desc_.change_max;
    P.init_shrink = desc_.init_shrink;
    P.stop_time = desc_.stop_time;
    const double *dtp = &ctl_->dt;
    engine_.set_halt(&ctl_->halt);
    int rc = CAMR_OK;
    for (int p = 0; p < 2 && !rc; ++p) {
        FillArgs fa;
        fa.sold_tab = p == 0 ? tab_b_ : tab_a_;  // parity 0: A is new, B old
        fa.owner = owner_;
        fa.ext = extv_;
        fa.has_ext = ext_ ? 1 : 0;
        fa.dom = dom_;
        fa.hc = hc_;
        fa.dtp = dtp;
        fa.g = desc_.const_grav;
        fa.add_grav = desc_.add_grav_src;
        PostArgs pa;
        pa.hc = hc_;
        pa.dtp = dtp;
        pa.g = desc_.const_grav;
        pa.add_grav = desc_.add_grav_src;
        pa.defer_post = 0;
        pa.dtmin = dtmin_;
        if (hipStreamBeginCapture(gstream_, hipStreamCaptureModeThreadLocal) != hipSuccess) {
            set_error("stream capture failed to start");
            rc = CAMR_ERR_HIP;
            break;
        }
        hipLaunchKernelGGL(k_step_dt, dim3(1), dim3(64), 0, gstream_, ctl_, dtmin_, status_, P);
        rc = engine_.fill(-1, p, fa, gstream_, status_, nullptr);
        if (!rc) rc = engine_.hydro(-1, p, true, nullptr, nullptr, 0, hc_, dtp, gstream_, status_,
                                    nullptr, &pa);
        hipGraph_t g = nullptr;
        const hipError_t ec = hipStreamEndCapture(gstream_, &g);
        if (!rc && (ec != hipSuccess || !g)) { set_error("stream capture failed"); rc = CAMR_ERR_HIP; }
        graph_[p] = g;
        if (!rc && hipGraphInstantiate(&gexec_[p], g, nullptr, nullptr, 0) != hipSuccess) {
            set_error("graph instantiation failed");
            rc = CAMR_ERR_HIP;
        }
    }
    engine_.set_halt(nullptr);
    if (rc) release_graphs();
    return rc;
}

// camr_level_run on the host: the reference dt sequence, one step at a time.
int HydroLevel::run(int nsteps, double *dts) {
    // profiled runs step on the host so that every launch is timed
    if (use_graph() && !prof_.enabled()) return run_graph(nsteps, dts);
    double dt = 0.0;
    for (int s = 0; s < nsteps; ++s) {
        int rc = (s == 0) ? initial_dt(&dt) : new_dt(dt, &dt);
        if (rc) return rc;
        if (dts) dts[s] = dt;
        rc = advance(dt);
        if (rc) return rc;
    }
    return sync();
}

// The same sequence with every step a replay of a captured graph: dt is
// computed on the device, so the host only queues replays.  A step that
// trips a density floor halts the remaining replays; the host then redoes
// that step exactly (resolve) and continues from there.
int HydroLevel::run_graph(int nsteps, double *dts) {
    if (nsteps <= 0) return sync();
    if (needs_ext_ && !ext_) {
        set_error("ghost cells outside every box of the level: set exterior data");
        return CAMR_ERR_SHAPE;
    }
    if (!ctl_) CAMR_HIP(hipMalloc(&ctl_, sizeof(StepCtl)));
    if (ctl_cap_ < nsteps) {
        if (ctl_dts_) (void)hipFree(ctl_dts_);
        ctl_dts_ = nullptr;
        CAMR_HIP(hipMalloc(&ctl_dts_, nsteps * sizeof(double)));
        ctl_cap_ = nsteps;
    }
    if (!gexec_[0]) {
        int rc = capture_graphs();
        if (rc) return rc;
    }
    int done = 0;
    double dt_old = 0.0;
    bool first = true;
    while (done < nsteps) {
        int rc = resolve();
        if (rc) return rc;
        if (!dt_ready_ && desc_.fixed_dt <= 0.0) {
            const unsigned long long init = (unsigned long long)0x7fefffffffffffffULL;
            launch_on(k_fill_u64, make_box1(), stream_, dtmin_, init);
            for (size_t b = 0; b < box_.size(); ++b)
                launch_on(k_estdt, box_[b], stream_, snew_[b], hc_.dx[0], hc_.dx[1], hc_.dx[2],
                          dtmin_, status_, ebf_);
            CAMR_HIP(hipGetLastError());
        }
        StepCtl c{};
        c.dt = dt_old;
        c.time = time_;
        c.first = first ? 1 : 0;
        c.halt = 0;
        c.nstep = 0;
        c.cap = ctl_cap_;
        c.dts = ctl_dts_;
        // from pageable memory: the copy completes before the call returns
        CAMR_HIP(hipMemcpyAsync(ctl_, &c, sizeof(c), hipMemcpyHostToDevice, stream_));
        CAMR_HIP(hipStreamSynchronize(stream_));
        const bool a0 = a_is_new_;
        const int todo = nsteps - done;
        for (int s = 0; s < todo; ++s) {
            a_is_new_ = !a_is_new_;  // swapTimeLevels
            CAMR_HIP(hipGraphLaunch(gexec_[parity()], stream_));
        }
        CAMR_HIP(hipMemcpyAsync(&c, ctl_, sizeof(c), hipMemcpyDeviceToHost, stream_));
        CAMR_HIP(hipStreamSynchronize(stream_));
        // the host buffer roles follow the steps actually taken
        a_is_new_ = (c.nstep & 1) ? !a0 : a0;
        snew_.clear();
        sold_.clear();
        for (size_t b = 0; b < box_.size(); ++b) {
            const Arr A = dense(bufA_[b], box_[b]), Bv = dense(bufB_[b], box_[b]);
            snew_.push_back(a_is_new_ ? A : Bv);
            sold_.push_back(a_is_new_ ? Bv : A);
        }
        if (c.nstep > 0 && dts)
            CAMR_HIP(hipMemcpy(dts + done, ctl_dts_, c.nstep * sizeof(double), hipMemcpyDeviceToHost));
        time_ = c.time;
        done += c.nstep;
        if (c.nstep > 0) {
            first = false;
            dt_old = c.dt;
            last_dt_ = c.dt;
            pending_ = true;   // resolve() redoes a floored last step
            dt_ready_ = true;  // the post pass left the CFL estimate in dtmin_
        }
        rc = check_status();
        if (rc) return rc;
        if (c.halt && c.nstep == 0) {
            rc = resolve();
            if (rc) return rc;
            if (!pending_) { set_error("captured step made no progress"); return CAMR_ERR_HIP; }
        }
    }
    return sync();
}

// ------------------------------------------------------------- adjoint
int HydroLevel::adjoint_scope() const {
    const char *why = nullptr;
    if (ebflags_) {
        // the embedded-boundary level (staged engine, PLM; set_eb refuses
        // gravity sources and the method of lines)
        bool nsw = false;
        for (int d = 0; d < 3; ++d)
            nsw |= desc_.hp.bclo[d] == BC_NOSLIPWALL || desc_.hp.bchi[d] == BC_NOSLIPWALL;
        if (desc_.hp.allow_small_energy != 1 || desc_.hp.allow_negative_energy != 0)
            why = "the adjoint needs allow_small_energy = 1, allow_negative_energy = 0";
        else if (nsw) why = "the embedded-boundary adjoint does not model NoSlipWall faces";
        else if (ext_ || needs_ext_) why = "the adjoint needs a level covering its domain";
        if (why) {
            set_error("%s", why);
            return CAMR_ERR_ADJOINT;
        }
        return CAMR_OK;
    }
    if (!fused_) why = "the adjoint runs on the fused engine";
    else if (desc_.hp.ppm_type != 1) why = "the adjoint differentiates the PPM path";
    else if (grav_src_values_) why = "the adjoint needs const_grav = 0";
    else if (desc_.hp.allow_small_energy != 1 || desc_.hp.allow_negative_energy != 0)
        why = "the adjoint needs allow_small_energy = 1, allow_negative_energy = 0";
    else if (!(desc_.is_periodic[0] && desc_.is_periodic[1] && desc_.is_periodic[2]))
        why = "the adjoint needs a periodic domain";
    else if (ext_ || needs_ext_) why = "the adjoint needs a level covering its domain";
    if (why) {
        set_error("%s", why);
        return CAMR_ERR_ADJOINT;
    }
    return CAMR_OK;
}

// The state of every box to / from a flat buffer (boxes one after another,
// state_off_ holding each box's offset), one launch over all boxes.
int HydroLevel::save_state(double *dst) {
    if (!state_off_) return CAMR_ERR_PARAM;
    launch_flat(kb_state_flat, (const DevFab *)(a_is_new_ ? tab_a_ : tab_b_), dst,
                (const int64_t *)state_off_, 1);
    return CAMR_OK;
}

int HydroLevel::load_state(const double *src) {
    int rc = resolve();
    if (rc) return rc;
    if (!state_off_) return CAMR_ERR_PARAM;
    launch_flat(kb_state_flat, (const DevFab *)(a_is_new_ ? tab_a_ : tab_b_),
                const_cast<double *>(src), (const int64_t *)state_off_, 0);
    dt_ready_ = false;
    return CAMR_OK;
}

int HydroLevel::adjoint_alloc(int nck, int every) {
    int rc = ebflags_ ? eb_adjoint_setup() : engine_.enable_adjoint();
    if (rc) return rc;
    state_doubles_ = 0;
    std::vector<int64_t> off;
    for (const Box &B : box_) {
        off.push_back((int64_t)state_doubles_);
        state_doubles_ += NVAR * B.npts();
    }
    if (!state_off_) {
        CAMR_HIP(hipMalloc(&state_off_, off.size() * sizeof(int64_t)));
        CAMR_HIP(hipMemcpy(state_off_, off.data(), off.size() * sizeof(int64_t), hipMemcpyHostToDevice));
    }
    if (abar_[0].empty()) {
        for (int q = 0; q < 2; ++q) {
            std::vector<DevFab> t(box_.size());
            for (size_t b = 0; b < box_.size(); ++b) {
                double *p = nullptr;
                CAMR_HIP(hipMalloc(&p, NVAR * box_[b].npts() * sizeof(double)));
                abar_[q].push_back(p);
                t[b].a = dense(p, box_[b]);
                t[b].b = box_[b];
            }
            CAMR_HIP(hipMalloc(&tab_abar_[q], box_.size() * sizeof(DevFab)));
            CAMR_HIP(hipMemcpy(tab_abar_[q], t.data(), box_.size() * sizeof(DevFab), hipMemcpyHostToDevice));
        }
    }
    while ((int)ckpt_.size() < nck) {
        double *p = nullptr;
        CAMR_HIP(hipMalloc(&p, state_doubles_ * sizeof(double)));
        ckpt_.push_back(p);
    }
    while ((int)blk_.size() < every) {
        double *p = nullptr;
        CAMR_HIP(hipMalloc(&p, state_doubles_ * sizeof(double)));
        blk_.push_back(p);
    }
    return CAMR_OK;
}

// From the level's state S_0: nsteps steps with frozen dts, checkpoints every
// `every` steps (AMReX-style restart copies kept on the device).
int HydroLevel::adjoint_forward(int nsteps, const double *dts, int every) {
    if (nsteps <= 0 || every <= 0 || !dts) { set_error("bad adjoint run arguments"); return CAMR_ERR_PARAM; }
    int rc = adjoint_scope();
    if (rc) return rc;
    rc = resolve();
    if (rc) return rc;
    const int nck = (nsteps + every - 1) / every;
    rc = adjoint_alloc(nck, every);
    if (rc) return rc;
    adj_dts_.assign(dts, dts + nsteps);
    adj_every_ = every;
    adj_time0_ = time_;
    // the states of the last block are kept as they are computed, so that
    // the first block of the backward pass need not recompute them
    const int tail = (nck - 1) * every;
    for (int s = 0; s < nsteps; ++s) {
        if (s % every == 0) {
            rc = resolve();
            if (!rc) rc = save_state(ckpt_[s / every]);
            if (rc) return rc;
        }
        rc = advance(dts[s]);
        if (!rc && s + 1 > tail && s + 1 < nsteps) {
            rc = resolve();
            if (!rc) rc = save_state(blk_[s + 1 - tail]);
        }
        if (rc) return rc;
    }
    adj_tail_ready_ = true;
    return sync();
}

// One step's adjoint: recompute the step from the current state S_s
// (fill + hydro up to the final fluxes; the post-pass adjoint applies the
// update to its cell itself),
// then run the adjoint kernels from abar_[acur_] (dL/dS_{s+1}) into the
// other set (dL/dS_s).  The level's state is S_s again on return.
int HydroLevel::adjoint_step(double dt, bool check) {
    if (ebflags_) return adjoint_step_eb(dt, check);
    int rc = resolve();
    if (rc) return rc;
    std::swap(snew_, sold_);
    a_is_new_ = !a_is_new_;
    set_dt_dev(dt);
    FillArgs fa;
    fa.sold_tab = a_is_new_ ? tab_b_ : tab_a_;
    fa.owner = owner_;
    fa.ext = extv_;
    fa.has_ext = 0;
    fa.dom = dom_;
    fa.hc = hc_;
    fa.dtp = dt_dev_;
    fa.g = desc_.const_grav;
    fa.add_grav = desc_.add_grav_src;
    rc = engine_.fill(-1, parity(), fa, stream_, status_, &prof_);
    // the update itself is applied in registers by the post-pass adjoint
    if (!rc) rc = engine_.hydro(-1, parity(), true, nullptr, nullptr, 0, hc_, dt_dev_, stream_,
                                status_, &prof_, nullptr, true);
    if (!rc) {
        const int out = 1 - acur_;
        // the adjoint scratch is all zero here: cleared before the first
        // step, and each kernel clears the adjoints it consumes
        AdjArgs a;
        a.sbar_in = tab_abar_[acur_];
        a.sbar_out = tab_abar_[out];
        a.sold_tab = fa.sold_tab;
        a.owner = owner_;
        a.dom = dom_;
        a.hc = hc_;
        a.dtp = dt_dev_;
        a.add_grav = desc_.add_grav_src;
        a.status = status_;
        rc = adjoint_step_kernels(engine_, parity(), a, stream_, &prof_);
        acur_ = out;
    }
    std::swap(snew_, sold_);
    a_is_new_ = !a_is_new_;
    dt_ready_ = false;
    if (rc) return rc;
    return check ? adjoint_flags() : CAMR_OK;
}

// The device flags after a stretch of asynchronous adjoint work: a density
// floor in a recomputed step (the forward step would have been redone with
// enforced floors, which the adjoint does not differentiate) or a device
// error of the adjoint kernels.
int HydroLevel::adjoint_flags() {
    CAMR_HIP(hipMemcpyAsync(h_status_, status_, sizeof(int), hipMemcpyDeviceToHost, stream_));
    CAMR_HIP(hipStreamSynchronize(stream_));
    if (*h_status_ & kFloorFlag) {
        *h_status_ &= ~kFloorFlag;
        CAMR_HIP(hipMemcpyAsync(status_, h_status_, sizeof(int), hipMemcpyHostToDevice, stream_));
        CAMR_HIP(hipStreamSynchronize(stream_));
        set_error("adjoint: the step takes a density floor, which is not differentiated");
        return CAMR_ERR_ADJOINT;
    }
    return status_error(*h_status_);
}

// The arrays of the embedded-boundary adjoint (allocated once) and the
// box tables (rebuilt on every call, so that geometry set later is seen).
int HydroLevel::eb_adjoint_setup() {
    const int nb = (int)box_.size();
    if (!eadj_fbuf_) {
        // regular geometry for the all-regular boxes: constant rows indexed
        // by i only (zero j / k / component strides)
        // (the flag box's x extent, plus the high nodes)
        const int64_t lo0 = ebf_.lo0, n0 = ebf_.js + 2;
        std::vector<double> ones((size_t)n0, 1.0), zeros((size_t)n0, 0.0);
        std::vector<int32_t> regf((size_t)n0, (int32_t)((1u << 27) - 1));
        double *po = nullptr, *pz = nullptr;
        int32_t *pf = nullptr;
        CAMR_HIP(hipMalloc(&po, n0 * sizeof(double)));
        CAMR_HIP(hipMalloc(&pz, n0 * sizeof(double)));
        CAMR_HIP(hipMalloc(&pf, n0 * sizeof(int32_t)));
        ebbuf_.push_back(po);
        ebbuf_.push_back(pz);
        ebbuf_.push_back(reinterpret_cast<double *>(pf));
        CAMR_HIP(hipMemcpy(po, ones.data(), n0 * sizeof(double), hipMemcpyHostToDevice));
        CAMR_HIP(hipMemcpy(pz, zeros.data(), n0 * sizeof(double), hipMemcpyHostToDevice));
        CAMR_HIP(hipMemcpy(pf, regf.data(), n0 * sizeof(int32_t), hipMemcpyHostToDevice));
        auto row = [&](double *p) {
            Arr a{};
            a.p = p;
            a.js = a.ks = a.ns = 0;
            a.lo0 = (int)lo0; a.lo1 = 0; a.lo2 = 0;
            return a;
        };
        unit_geom_.f.p = pf;
        unit_geom_.f.lo0 = (int)lo0; unit_geom_.f.lo1 = 0; unit_geom_.f.lo2 = 0;
        unit_geom_.f.js = 0; unit_geom_.f.ks = 0;
        unit_geom_.vf = row(po);
        for (int d = 0; d < 3; ++d) { unit_geom_.ap[d] = row(po); unit_geom_.fc[d] = row(pz); }
        // the boxes of the recompute: every box with cut cells or regular
        // alone (its level index), then the virtual boxes (-1 - index)
        eadj_ids_.clear();
        for (int b = 0; b < nb; ++b)
            if (kind_[b] != 2 && (vof_.empty() || vof_[b] < 0)) eadj_ids_.push_back(b);
        for (size_t v = 0; v < vbox_.size(); ++v) eadj_ids_.push_back(-1 - (int)v);
        auto r32 = [](size_t n) { return (n + 31) / 32 * 32; };
        size_t fn = 0, an = 0;
        for (int id : eadj_ids_) {
            const Box B = id >= 0 ? box_[id] : vbox_[-1 - id], G = grow(B, ng_);
            fn += r32((size_t)(QVAR + NQAUX) * G.npts());
            for (int d = 0; d < 3; ++d) fn += r32((size_t)NVAR * nodes(B, d).npts());
            fn += r32(umdrv_eb_scratch_doubles(B, G, true));
            an += r32((size_t)(QVAR + NQAUX) * G.npts());
            an += r32(umdrv_eb_scratch_doubles(B, G, false));
        }
        for (int b = 0; b < nb; ++b)
            an += r32((size_t)NVAR * box_[b].npts()) + r32((size_t)NVAR * grow(box_[b], ng_).npts());
        for (const Box &V : vbox_)
            an += r32((size_t)NVAR * V.npts()) + r32((size_t)NVAR * grow(V, ng_).npts());
        if (!vbox_.empty()) {
            CAMR_HIP(hipMalloc(&vga_dev_, vbox_.size() * sizeof(VGather)));
            ebbuf_.push_back(reinterpret_cast<double *>(vga_dev_));
        }
        CAMR_HIP(hipMalloc(&eadj_fbuf_, fn * sizeof(double)));
        ebbuf_.push_back(eadj_fbuf_);
        CAMR_HIP(hipMalloc(&eadj_abuf_, an * sizeof(double)));
        ebbuf_.push_back(eadj_abuf_);
        eadj_an_ = an;

        const size_t n = eadj_ids_.size();
        CAMR_HIP(hipMalloc(&eadj_ftab_, std::max<size_t>(n, 1) * sizeof(EbBox)));
        ebbuf_.push_back(reinterpret_cast<double *>(eadj_ftab_));
        CAMR_HIP(hipMalloc(&eadj_tab_, std::max<size_t>(n, 1) * sizeof(EbAdjEnt)));
        ebbuf_.push_back(reinterpret_cast<double *>(eadj_tab_));
        CAMR_HIP(hipMalloc(&tab_ub_, nb * sizeof(DevFab)));
        ebbuf_.push_back(reinterpret_cast<double *>(tab_ub_));
        CAMR_HIP(hipMalloc(&tab_sbb_, nb * sizeof(DevFab)));
        ebbuf_.push_back(reinterpret_cast<double *>(tab_sbb_));
        CAMR_HIP(hipMalloc(&geom_dev_, nb * sizeof(EbGeom)));
        ebbuf_.push_back(reinterpret_cast<double *>(geom_dev_));
        CAMR_HIP(hipMalloc(&kinds_dev_, nb * sizeof(int)));
        ebbuf_.push_back(reinterpret_cast<double *>(kinds_dev_));
    }
    // tables
    auto r32 = [](size_t n) { return (n + 31) / 32 * 32; };
    double *fp = eadj_fbuf_, *ap = eadj_abuf_;
    std::vector<Arr> ubv(box_.size()), sbbv(box_.size());
    std::vector<EbGeom> geo(box_.size());
    eadj_fh_.assign(eadj_ids_.size(), EbBox{});
    eadj_h_.assign(eadj_ids_.size(), EbAdjEnt{});
    // the state-sized adjoint arrays of every box come first
    for (size_t b = 0; b < box_.size(); ++b) {
        const Box B = box_[b], G = grow(B, ng_);
        ubv[b] = dense(ap, B);
        ap += r32((size_t)NVAR * B.npts());
        sbbv[b] = dense(ap, G);
        ap += r32((size_t)NVAR * G.npts());
        geo[b] = kind_[b] == 1 ? ebg_[b] : unit_geom_;
        if (kind_[b] == 1 && !ebg_[b].vf.p) {
            set_error("box %d has cut cells but no cut-cell data (camr_level_set_box_eb)", (int)b);
            return CAMR_ERR_PARAM;
        }
    }
    // each virtual box's dL/dS_new (the rows of its boxes view it) and
    // dL/dSborder over grow(v, 6), added into its boxes' after the stages
    std::vector<Arr> vub(vbox_.size());
    std::vector<VGather> vga = vgh_;
    for (size_t v = 0; v < vbox_.size(); ++v) {
        const Box V = vbox_[v], G = grow(V, ng_);
        vub[v] = dense(ap, V);
        ap += r32((size_t)NVAR * V.npts());
        vga[v].sb = dense(ap, G);
        ap += r32((size_t)NVAR * G.npts());
    }
    for (size_t b = 0; b < box_.size(); ++b)
        if (!vof_.empty() && vof_[b] >= 0) ubv[b] = vub[vof_[b]];
    std::vector<std::pair<size_t, size_t>> skip;  // face-state buffers, in address order
    for (size_t c = 0; c < eadj_ids_.size(); ++c) {
        const int id = eadj_ids_[c], v = -1 - id;
        const Box B = id >= 0 ? box_[id] : vbox_[v], G = grow(B, ng_);
        UmdrvEbArgs A;
        A.bx = B;
        A.qbox = G;
        A.hc = hc_;
        for (int d = 0; d < 3; ++d) A.periodic[d] = desc_.is_periodic[d];
        A.dtp = dt_dev_;
        A.uin = id >= 0 ? sbv_[id] : vgh_[v].sb;
        A.dsdt = id >= 0 ? dsv_[id] : vds_[v];
        A.q = dense(fp, G);
        A.qaux = dense(fp + (size_t)QVAR * G.npts(), G);
        A.srcq = A.q;  // no sources
        A.has_src = 0;
        fp += r32((size_t)(QVAR + NQAUX) * G.npts());
        for (int d = 0; d < 3; ++d) {
            A.flx[d] = dense(fp, nodes(B, d));
            fp += r32((size_t)NVAR * nodes(B, d).npts());
        }
        A.g = id >= 0 ? geo[id] : unit_geom_;
        A.regular = (id < 0 || kind_[id] == 0) ? 1 : 0;
        A.noflx = 1;  // the adjoint reads no output flux
        eb_box_setup(A, fp, &eadj_fh_[c], true);
        fp += r32(umdrv_eb_scratch_doubles(B, G, true));
        UmdrvEbArgs Ab = A;
        Ab.q = dense(ap, G);
        Ab.qaux = dense(ap + (size_t)QVAR * G.npts(), G);
        Ab.srcq = Ab.q;
        ap += r32((size_t)(QVAR + NQAUX) * G.npts());
        Ab.uin = id >= 0 ? sbbv[id] : vga[v].sb;
        Ab.dsdt = id >= 0 ? ubv[id] : vub[v];
        // the adjoint side shares one buffer of face states among the pairs
        // and one among the final solves (set, not accumulated)
        eb_box_setup(Ab, ap, &eadj_h_[c].A, false);
        ap += r32(umdrv_eb_scratch_doubles(B, G, false));
        eadj_h_[c].F = eadj_fh_[c];
        const EbBox &Ad = eadj_h_[c].A;
        const size_t x0 = (size_t)(Ad.qmT[0].p - eadj_abuf_);
        const size_t x1 = (size_t)(Ad.qpF[0].p - eadj_abuf_) +
                          (size_t)NS * layout_of(B, G).bxg1.npts();
        skip.push_back({x0, x1});
    }
    {
        // clear the complement of the face-state buffers
        eadj_zs_.clear();
        eadj_zl_.clear();
        size_t at = 0;
        for (const auto &x : skip) {
            if (x.first > at) { eadj_zs_.push_back(at); eadj_zl_.push_back(x.first - at); }
            at = x.second;
        }
        if (eadj_an_ > at) { eadj_zs_.push_back(at); eadj_zl_.push_back(eadj_an_ - at); }
    }
    std::vector<DevFab> tu(box_.size()), ts(box_.size());
    std::vector<int> kd(box_.size());
    for (size_t b = 0; b < box_.size(); ++b) {
        tu[b].a = ubv[b]; tu[b].b = box_[b];
        ts[b].a = sbbv[b]; ts[b].b = box_[b];
        kd[b] = kind_[b];
    }
    const size_t n = eadj_ids_.size();
    if (n) {
        CAMR_HIP(hipMemcpy(eadj_ftab_, eadj_fh_.data(), n * sizeof(EbBox), hipMemcpyHostToDevice));
        CAMR_HIP(hipMemcpy(eadj_tab_, eadj_h_.data(), n * sizeof(EbAdjEnt), hipMemcpyHostToDevice));
    }
    CAMR_HIP(hipMemcpy(tab_ub_, tu.data(), tu.size() * sizeof(DevFab), hipMemcpyHostToDevice));
    CAMR_HIP(hipMemcpy(tab_sbb_, ts.data(), ts.size() * sizeof(DevFab), hipMemcpyHostToDevice));
    if (!vga.empty())
        CAMR_HIP(hipMemcpy(vga_dev_, vga.data(), vga.size() * sizeof(VGather), hipMemcpyHostToDevice));
    if (!eadj_cells_ && !cut_cells_h_.empty()) {
        // the cut cells with the adjoint table's entries
        std::vector<int> pos(box_.size(), -1);
        for (size_t c = 0; c < eadj_ids_.size(); ++c)
            if (eadj_ids_[c] >= 0) pos[eadj_ids_[c]] = (int)c;
        std::vector<CutCell> cells = cut_cells_h_;
        for (CutCell &x : cells) x.e = pos[cut_ids_[x.e]];
        neadj_cells_ = (int)cells.size();
        CAMR_HIP(hipMalloc(&eadj_cells_, cells.size() * sizeof(CutCell)));
        ebbuf_.push_back(reinterpret_cast<double *>(eadj_cells_));
        CAMR_HIP(hipMemcpy(eadj_cells_, cells.data(), cells.size() * sizeof(CutCell),
                           hipMemcpyHostToDevice));
    }
    CAMR_HIP(hipMemcpy(geom_dev_, geo.data(), geo.size() * sizeof(EbGeom), hipMemcpyHostToDevice));
    CAMR_HIP(hipMemcpy(kinds_dev_, kd.data(), kd.size() * sizeof(int), hipMemcpyHostToDevice));
    return CAMR_OK;
}

// One step's adjoint on an embedded-boundary level: recompute the step from
// S_s (clean_state, FillPatch, primitives, the cut-cell stages of every box
// that is not covered), then the adjoint kernels from abar_[acur_]
// (dL/dS_{s+1}) into the other set (dL/dS_s).
int HydroLevel::adjoint_step_eb(double dt, bool check) {
    int rc = resolve();
    if (rc) return rc;
    std::swap(snew_, sold_);
    a_is_new_ = !a_is_new_;
    set_dt_dev(dt);
    const DevFab *told = a_is_new_ ? tab_b_ : tab_a_;
    clean_all(told, true);
    expand_all(told, true);
    const int64_t vn = vbox_points();
    for (size_t b = 0; b < box_.size(); ++b)  // covered boxes: no hydro, dS/dt = 0
        if (kind_[b] == 2)
            launch_on(k_fill_value, box_[b], stream_, dsv_[b], (int)NVAR, 0.0);
    const int n = (int)eadj_ids_.size();
    rc = eb_prim_batch(eadj_ftab_, eadj_fh_.data(), n, hc_, status_, stream_, &prof_);
    if (!rc)
        rc = umdrv_eb_batch(eadj_ftab_, eadj_fh_.data(), n, hc_, dt_dev_, 0, stream_, &prof_,
                            eadj_cells_, neadj_cells_);
    if (!rc) {
        const int out = 1 - acur_;
        zero_ranges(eadj_abuf_, eadj_zs_, eadj_zl_, stream_);
        launch_all(kb_zero_state, 0, (const DevFab *)tab_abar_[out]);
        launch_all(kb_ebadj_post, 0, (const DevFab *)tab_sb_, (const DevFab *)tab_dsdt_,
                   (const DevFab *)tab_abar_[acur_], (const DevFab *)tab_ub_,
                   (const DevFab *)tab_sbb_, (const double *)dt_dev_, hc_,
                   desc_.add_grav_src, status_, ebf_);
        rc = eb_adjoint_batch(eadj_tab_, eadj_h_.data(), n, hc_, dt_dev_, stream_, &prof_,
                              eadj_cells_, neadj_cells_);
        if (vn) {
            FlatBlocks fb;
            const dim3 grid = flat_blocks(
                (int)vbox_.size(),
                [&](int v) -> int64_t { return grid_for(grow(vbox_[v], ng_).npts()); }, fb);
            hipLaunchKernelGGL(kb_vscatter, grid, dim3(kBlock), 0, stream_,
                               (const VGather *)vga_dev_, (const int *)vg_ids_,
                               (const DevFab *)tab_sbb_, ng_, kEbAdjComps, fb);
        }
        launch_all(kb_ebadj_fill, ng_, (const DevFab *)tab_sbb_, ng_, told,
                   (const DevFab *)tab_abar_[out], (const int *)owner_, dom_, hc_, status_);
        launch_all(kb_ebadj_old, 0, (const DevFab *)tab_abar_[out], told, hc_, status_);
        acur_ = out;
    }
    std::swap(snew_, sold_);
    a_is_new_ = !a_is_new_;
    dt_ready_ = false;
    if (rc) { set_error("kernel launch failed"); return rc; }
    return check ? adjoint_flags() : CAMR_OK;
}

// abar_[acur_] = d(w . F)/dS of the force on the embedded body (eb_force)
// at the level's current state: the seed of the force's adjoint.
int HydroLevel::eb_force_seed(const double w[3]) {
    if (!ebflags_) { set_error("the level has no embedded boundary"); return CAMR_ERR_PARAM; }
    if (abar_[0].empty() || !geom_dev_) { set_error("no adjoint forward run"); return CAMR_ERR_PARAM; }
    int rc = resolve();
    if (rc) return rc;
    const double dx2 = desc_.hp.dx[0] * desc_.hp.dx[0];
    launch_all(kb_ebadj_force, 0, (const DevFab *)(a_is_new_ ? tab_a_ : tab_b_),
               (const DevFab *)tab_abar_[acur_], (const EbGeom *)geom_dev_,
               (const int *)kinds_dev_, w[0], w[1], w[2], dx2, hc_);
    CAMR_HIP(hipGetLastError());
    return sync();
}

int HydroLevel::adjoint_backward() {
    const int nsteps = (int)adj_dts_.size();
    if (nsteps == 0 || abar_[0].empty()) { set_error("no adjoint forward run"); return CAMR_ERR_PARAM; }
    const int every = adj_every_;
    const int nck = (nsteps + every - 1) / every;
    int rc = resolve();
    if (rc) return rc;
    if (!ebflags_) zero_doubles(engine_.adjoint_base(), engine_.scratch_total(), stream_);
    for (int c = nck - 1; c >= 0 && !rc; --c) {
        const int start = c * every;
        const int len = std::min(every, nsteps - start);
        rc = load_state(ckpt_[c]);
        time_ = adj_time0_;
        // states S_start .. S_start+len-1 of this block (those of the last
        // block are still there from the forward run).  No host check per
        // step: a floor would have refused the run, and the block's flags
        // are read once at its end (they accumulate on the device).
        const bool kept = c == nck - 1 && adj_tail_ready_;
        adj_tail_ready_ = false;
        for (int k = 1; k < len && !rc && !kept; ++k) {
            rc = advance(adj_dts_[start + k - 1]);
            pending_ = false;
            if (!rc) rc = save_state(blk_[k]);
        }
        for (int k = len - 1; k >= 0 && !rc; --k) {
            rc = load_state(k == 0 ? ckpt_[c] : blk_[k]);
            if (!rc) rc = adjoint_step(adj_dts_[start + k], false);
        }
        if (!rc) rc = adjoint_flags();
    }
    if (rc) return rc;
    rc = load_state(ckpt_[0]);
    time_ = adj_time0_;
    if (rc) return rc;
    return sync();
}

int HydroLevel::adj_upload(int b, const double *h) {
    if (b < 0 || b >= (int)box_.size() || abar_[0].empty()) { set_error("bad box index or no adjoint run"); return CAMR_ERR_PARAM; }
    CAMR_HIP(hipStreamSynchronize(stream_));
    CAMR_HIP(hipMemcpy(abar_[acur_][b], h, NVAR * box_[b].npts() * sizeof(double), hipMemcpyHostToDevice));
    return CAMR_OK;
}

int HydroLevel::adj_download(int b, double *h) {
    if (b < 0 || b >= (int)box_.size() || abar_[0].empty()) { set_error("bad box index or no adjoint run"); return CAMR_ERR_PARAM; }
    CAMR_HIP(hipStreamSynchronize(stream_));
    CAMR_HIP(hipMemcpy(h, abar_[acur_][b], NVAR * box_[b].npts() * sizeof(double), hipMemcpyDeviceToHost));
    return CAMR_OK;
}

// Settle a fused step: wait for it and, if a density floor was flagged,
// recompute it with the exact pass-by-pass sequence.
int HydroLevel::resolve() {
    if (!pending_) return CAMR_OK;
    pending_ = false;
    CAMR_HIP(hipMemcpyAsync(h_status_, status_, sizeof(int), hipMemcpyDeviceToHost, stream_));
    CAMR_HIP(hipStreamSynchronize(stream_));
    if (*h_status_ & kFloorFlag) {
        *h_status_ &= ~kFloorFlag;
        CAMR_HIP(hipMemcpyAsync(status_, h_status_, sizeof(int), hipMemcpyHostToDevice, stream_));
        set_dt_dev(last_dt_);  // graph replays keep dt on the device only
        int rc = advance_exact(last_dt_);
        dt_ready_ = false;
        if (rc) return rc;
        // the floor bit only gates enforce_min_density within a step: the
        // redone step leaves it set, the next step starts without it
        CAMR_HIP(hipMemcpyAsync(h_status_, status_, sizeof(int), hipMemcpyDeviceToHost, stream_));
        CAMR_HIP(hipStreamSynchronize(stream_));
        if (*h_status_ & kFloorFlag) {
            *h_status_ &= ~kFloorFlag;
            CAMR_HIP(hipMemcpyAsync(status_, h_status_, sizeof(int), hipMemcpyHostToDevice, stream_));
            CAMR_HIP(hipStreamSynchronize(stream_));
        }
    }
    return check_status();
}

// estTimeStep (CAMR.cpp:490-563)
int HydroLevel::est_time_step(double *est) {
    if (desc_.fixed_dt > 0.0) { *est = desc_.fixed_dt; return CAMR_OK; }
    double estdt = desc_.max_dt;
    const double max_dt_over_cfl = desc_.max_dt / desc_.cfl;
    double estdt_hydro = max_dt_over_cfl;
    int rc0 = resolve();
    if (rc0) return rc0;
    if (!dt_ready_) {
        const unsigned long long init = (unsigned long long)0x7fefffffffffffffULL;  // DBL_MAX
        launch_on(k_fill_u64, make_box1(), stream_, dtmin_, init);
        launch_all(kb_estdt, 0, (const DevFab *)(a_is_new_ ? tab_a_ : tab_b_), hc_.dx[0],
                   hc_.dx[1], hc_.dx[2], dtmin_, status_, ebf_);
        CAMR_HIP(hipGetLastError());
        dt_ready_ = true;
    }
    CAMR_HIP(hipMemcpyAsync(h_dtmin_, dtmin_, sizeof(unsigned long long), hipMemcpyDeviceToHost, stream_));
    int rc = sync();
    if (rc) return rc;
    double dt;
    unsigned long long bits = *h_dtmin_;
    memcpy(&dt, &bits, sizeof(dt));
    estdt_hydro = vmin(estdt_hydro, dt);
    estdt_hydro *= desc_.cfl;
    if (estdt_hydro < estdt) estdt = estdt_hydro;
    *est = estdt;
    return CAMR_OK;
}

// computeInitialDt (CAMR.cpp:638-676) with initialTimeStep (:474-488)
int HydroLevel::initial_dt(double *dt) {
    double est;
    int rc = est_time_step(&est);
    if (rc) return rc;
    double dt_0 = DBL_MAX;
    dt_0 = vmin(dt_0, 1 * (desc_.init_shrink * est));
    const double dt_eps = 0.001 * dt_0;
    if (desc_.stop_time >= 0.0 && (time_ + dt_0) > (desc_.stop_time - dt_eps))
        dt_0 = desc_.stop_time - time_;
    *dt = dt_0 / 1;
    return CAMR_OK;
}

// computeNewDt (CAMR.cpp:566-635), single level, no regrid
int HydroLevel::new_dt(double dt_old, double *dt) {
    double dt_min;
    int rc = est_time_step(&dt_min);
    if (rc) return rc;
    if (desc_.fixed_dt <= 0.0) dt_min = vmin(dt_min, desc_.change_max * dt_old);
    double dt_0 = DBL_MAX;
    dt_0 = vmin(dt_0, 1 * dt_min);
    const double dt_eps = 0.001 * dt_0;
    if (desc_.stop_time >= 0.0 && (time_ + dt_0) > (desc_.stop_time - dt_eps))
        dt_0 = desc_.stop_time - time_;
    *dt = dt_0 / 1;
    return CAMR_OK;
}

} // namespace camr

// ============================================ coarse-fine synchronisation
// The flux register between two levels (AMReX's YAFluxRegister as CAMR uses
// it, CAMR.H:310), average_down (CAMR.cpp:1047-1066) and the two-level
// subcycled step (Amr::timeStep with CAMR::post_timestep, CAMR.cpp:
// 679-699).  Coarse boundary cells that border two different fine boxes
// are not supported (AMReX weighs its overlapping coarse-fine patches
// there); the register refuses such layouts.
namespace camr {

static double cell_volume(const double *dx) {
    double v = dx[0];
    v *= dx[1];
    v *= dx[2];
    return v;
}

// periodic image of a coarse index; false outside a non-periodic domain
static bool host_reg_map(const DomInfo &D, const int *c, int *m) {
    for (int d = 0; d < 3; ++d) {
        const int lo = D.domlo[d], n = D.domhi[d] - lo + 1;
        int x = c[d];
        if (x < lo || x > lo + n - 1) {
            if (!D.periodic[d]) return false;
            x = lo + ((x - lo) % n + n) % n;
        }
        m[d] = x;
    }
    return true;
}

static int check_refines(const camr_level_desc &f, const camr_level_desc &c, int ratio) {
    if (ratio < 2) { set_error("refinement ratio must be at least 2"); return CAMR_ERR_PARAM; }
    for (int d = 0; d < 3; ++d)
        if (f.hp.domlo[d] != c.hp.domlo[d] * ratio ||
            f.hp.domhi[d] + 1 != (c.hp.domhi[d] + 1) * ratio) {
            set_error("the fine domain is not the coarse domain refined by %d", ratio);
            return CAMR_ERR_SHAPE;
        }
    return CAMR_OK;
}

int flux_register_create(HydroLevel &C, HydroLevel &F, int ratio, FluxReg **out) {
    int rc = check_refines(F.desc_, C.desc_, ratio);
    if (rc) return rc;
    if (C.ebflags_ || F.ebflags_) { set_error("flux registers with embedded boundaries are not built"); return CAMR_ERR_PARAM; }
    if (C.reg_crse_ || F.reg_fine_) { set_error("a flux register is already attached"); return CAMR_ERR_PARAM; }
    if (C.desc_.hp.do_mol != F.desc_.hp.do_mol) { set_error("both levels need the same integrator"); return CAMR_ERR_PARAM; }
    for (const Box &B : F.box_)
        for (int d = 0; d < 3; ++d)
            if (B.lo[d] % ratio || (B.hi[d] + 1) % ratio) {
                set_error("fine boxes must be aligned to the refinement ratio");
                return CAMR_ERR_SHAPE;
            }
    const DomInfo &D = C.dom_;
    const int64_t n0 = D.domhi[0] - D.domlo[0] + 1, n1 = D.domhi[1] - D.domlo[1] + 1,
                  n2 = D.domhi[2] - D.domlo[2] + 1;
    auto idx = [&](const int *m) {
        return (int64_t)(m[0] - D.domlo[0]) + (int64_t)(m[1] - D.domlo[1]) * n0 +
               (int64_t)(m[2] - D.domlo[2]) * n0 * n1;
    };
    std::vector<uint8_t> flag((size_t)(n0 * n1 * n2), 0);
    std::vector<int> cov(flag.size(), -1);
    for (size_t fb = 0; fb < F.box_.size(); ++fb) {
        const Box &B = F.box_[fb];
        for (int k = B.lo[2] / ratio; k <= B.hi[2] / ratio; ++k)
            for (int j = B.lo[1] / ratio; j <= B.hi[1] / ratio; ++j)
                for (int i = B.lo[0] / ratio; i <= B.hi[0] / ratio; ++i) {
                    const int c[3] = {i, j, k};
                    flag[idx(c)] = kRegFine;
                    cov[idx(c)] = (int)fb;
                }
    }
    for (const Box &B : F.box_)
        for (int k = B.lo[2] / ratio - 1; k <= B.hi[2] / ratio + 1; ++k)
            for (int j = B.lo[1] / ratio - 1; j <= B.hi[1] / ratio + 1; ++j)
                for (int i = B.lo[0] / ratio - 1; i <= B.hi[0] / ratio + 1; ++i) {
                    const int c[3] = {i, j, k};
                    int m[3];
                    if (host_reg_map(D, c, m) && flag[idx(m)] != kRegFine) flag[idx(m)] = kRegBnd;
                }
    for (int k = D.domlo[2]; k <= D.domhi[2]; ++k)
        for (int j = D.domlo[1]; j <= D.domhi[1]; ++j)
            for (int i = D.domlo[0]; i <= D.domhi[0]; ++i) {
                const int c[3] = {i, j, k};
                if (flag[idx(c)] != kRegBnd) continue;
                int seen = -1;
                for (int d = 0; d < 3; ++d)
                    for (int sg = -1; sg <= 1; sg += 2) {
                        int nb[3] = {i, j, k}, m[3];
                        nb[d] += sg;
                        if (!host_reg_map(D, nb, m)) continue;
                        const int fb = cov[idx(m)];
                        if (fb < 0) continue;
                        if (seen >= 0 && fb != seen) {
                            set_error("coarse cell (%d,%d,%d) borders two fine boxes", i, j, k);
                            return CAMR_ERR_SHAPE;
                        }
                        seen = fb;
                    }
            }
    // which slabs of which fine box border coarse boundary cells
    std::vector<int> slabs(F.box_.size(), 0);
    for (size_t fb = 0; fb < F.box_.size(); ++fb) {
        Box cb = F.box_[fb];
        for (int d = 0; d < 3; ++d) { cb.lo[d] /= ratio; cb.hi[d] = (cb.hi[d] + 1) / ratio - 1; }
        for (int d = 0; d < 3; ++d)
            for (int side = 0; side < 2; ++side) {
                Box sl = cb;
                sl.lo[d] = sl.hi[d] = side ? cb.hi[d] + 1 : cb.lo[d] - 1;
                bool any = false;
                for (int k = sl.lo[2]; k <= sl.hi[2] && !any; ++k)
                    for (int j = sl.lo[1]; j <= sl.hi[1] && !any; ++j)
                        for (int i = sl.lo[0]; i <= sl.hi[0] && !any; ++i) {
                            const int c[3] = {i, j, k};
                            int m[3];
                            any = host_reg_map(D, c, m) && flag[idx(m)] == kRegBnd;
                        }
                if (any) slabs[fb] |= 1 << (2 * d + side);
            }
    }
    if ((rc = C.resolve()) || (rc = F.resolve())) return rc;
    CAMR_HIP(hipStreamSynchronize(C.stream_));
    CAMR_HIP(hipStreamSynchronize(F.stream_));
    std::unique_ptr<FluxReg> R(new FluxReg());
    R->slabs = slabs;
    R->C = &C;
    R->F = &F;
    R->ratio = ratio;
    R->dx1c = cell_volume(C.desc_.hp.dx);
    R->dx1f = cell_volume(F.desc_.hp.dx);
    CAMR_HIP(hipMalloc(&R->flag, flag.size()));
    CAMR_HIP(hipMemcpy(R->flag, flag.data(), flag.size(), hipMemcpyHostToDevice));
    const size_t nb = C.box_.size();
    std::vector<DevFab> tc(nb), tf(nb);
    R->crse.assign(nb, nullptr);
    R->cfp.assign(nb, nullptr);
    int64_t half = 0;
    for (size_t b = 0; b < nb; ++b) half += (int64_t)NVAR * C.box_[b].npts();
    R->pool_n = 2 * half;
    CAMR_HIP(hipMalloc(&R->pool, R->pool_n * sizeof(double)));
    CAMR_HIP(hipMemset(R->pool, 0, R->pool_n * sizeof(double)));
    int64_t off = 0;
    for (size_t b = 0; b < nb; ++b) {
        R->crse[b] = R->pool + off;
        R->cfp[b] = R->pool + half + off;
        off += (int64_t)NVAR * C.box_[b].npts();
        tc[b].a = dense(R->crse[b], C.box_[b]);
        tc[b].b = C.box_[b];
        tf[b].a = dense(R->cfp[b], C.box_[b]);
        tf[b].b = C.box_[b];
    }
    CAMR_HIP(hipMalloc(&R->tab_crse, nb * sizeof(DevFab)));
    CAMR_HIP(hipMalloc(&R->tab_cfp, nb * sizeof(DevFab)));
    CAMR_HIP(hipMemcpy(R->tab_crse, tc.data(), nb * sizeof(DevFab), hipMemcpyHostToDevice));
    CAMR_HIP(hipMemcpy(R->tab_cfp, tf.data(), nb * sizeof(DevFab), hipMemcpyHostToDevice));
    C.release_graphs();
    F.release_graphs();
    C.reg_crse_ = R.get();
    F.reg_fine_ = R.get();
    *out = R.release();
    return CAMR_OK;
}

// CrseAdd / FineAdd of box b's fluxes, on the stream that computed them,
// with fac_for_reflux * dt (CAMR_construct_hydro_source.cpp:24)
void HydroLevel::reg_add(size_t b, const Arr *flx, hipStream_t s) {
    const double fac = desc_.hp.do_mol ? 0.5 : 1.0;
    if (reg_crse_) {
        FluxReg &R = *reg_crse_;
        launch_on(k_reg_crse_add, box_[b], s, dense(R.crse[b], box_[b]), flx[0], flx[1], flx[2],
                  (const uint8_t *)R.flag, dom_, (const double *)dt_dev_, fac, R.dx1c);
    }
    if (reg_fine_) {
        FluxReg &R = *reg_fine_;
        const int r = R.ratio;
        Box cb = box_[b];
        for (int d = 0; d < 3; ++d) { cb.lo[d] /= r; cb.hi[d] = (cb.hi[d] + 1) / r - 1; }
        for (int d = 0; d < 3; ++d)
            for (int side = 0; side < 2; ++side) {
                if (!(R.slabs[b] >> (2 * d + side) & 1)) continue;
                Box sl = cb;
                sl.lo[d] = sl.hi[d] = side ? cb.hi[d] + 1 : cb.lo[d] - 1;
                if (d == 0)
                    launch_on(k_reg_fine_add<0>, sl, s, (const DevFab *)R.tab_cfp, (const int *)R.C->owner_,
                              R.C->dom_, (const uint8_t *)R.flag, flx[0], side, r,
                              (const double *)dt_dev_, fac, R.dx1f);
                else if (d == 1)
                    launch_on(k_reg_fine_add<1>, sl, s, (const DevFab *)R.tab_cfp, (const int *)R.C->owner_,
                              R.C->dom_, (const uint8_t *)R.flag, flx[1], side, r,
                              (const double *)dt_dev_, fac, R.dx1f);
                else
                    launch_on(k_reg_fine_add<2>, sl, s, (const DevFab *)R.tab_cfp, (const int *)R.C->owner_,
                              R.C->dom_, (const uint8_t *)R.flag, flx[2], side, r,
                              (const double *)dt_dev_, fac, R.dx1f);
            }
    }
}

// The fused engine's adds, each one launch over a table of records built
// on first use (the canonical flux arrays and the register never move).
int HydroLevel::reg_add_fused() {
    const double fac = desc_.hp.do_mol ? 0.5 : 1.0;
    if (reg_crse_) {
        FluxReg &R = *reg_crse_;
        if (!R.crse_recs) {
            std::vector<RegCrseRec> v(box_.size());
            for (size_t b = 0; b < box_.size(); ++b) {
                v[b].r = range_of(box_[b]);
                v[b].d = dense(R.crse[b], box_[b]);
                for (int d = 0; d < 3; ++d) v[b].f[d] = engine_.flux_view((int)b, d);
                R.crse_pts = std::max<int64_t>(R.crse_pts, box_[b].npts());
            }
            CAMR_HIP(hipMalloc(&R.crse_recs, v.size() * sizeof(RegCrseRec)));
            CAMR_HIP(hipMemcpy(R.crse_recs, v.data(), v.size() * sizeof(RegCrseRec), hipMemcpyHostToDevice));
            R.ncrse_recs = (int)v.size();
        }
        hipLaunchKernelGGL(kb_reg_crse_add, dim3(grid_for(R.crse_pts), R.ncrse_recs), dim3(kBlock), 0,
                           stream_, (const RegCrseRec *)R.crse_recs, (const uint8_t *)R.flag, dom_,
                           (const double *)dt_dev_, fac, R.dx1c);
    }
    if (reg_fine_) {
        FluxReg &R = *reg_fine_;
        const int r = R.ratio;
        if (!R.fine_recs) {
            std::vector<RegFineRec> v;
            for (size_t b = 0; b < box_.size(); ++b) {
                Box cb = box_[b];
                for (int d = 0; d < 3; ++d) { cb.lo[d] /= r; cb.hi[d] = (cb.hi[d] + 1) / r - 1; }
                for (int d = 0; d < 3; ++d)
                    for (int side = 0; side < 2; ++side) {
                        if (!(R.slabs[b] >> (2 * d + side) & 1)) continue;
                        Box sl = cb;
                        sl.lo[d] = sl.hi[d] = side ? cb.hi[d] + 1 : cb.lo[d] - 1;
                        RegFineRec e;
                        e.r = range_of(sl);
                        e.f = engine_.flux_view((int)b, d);
                        e.d = d;
                        e.side = side;
                        v.push_back(e);
                        R.fine_pts = std::max<int64_t>(R.fine_pts, sl.npts());
                    }
            }
            if (v.empty()) return CAMR_OK;
            CAMR_HIP(hipMalloc(&R.fine_recs, v.size() * sizeof(RegFineRec)));
            CAMR_HIP(hipMemcpy(R.fine_recs, v.data(), v.size() * sizeof(RegFineRec), hipMemcpyHostToDevice));
            R.nfine_recs = (int)v.size();
        }
        hipLaunchKernelGGL(kb_reg_fine_add, dim3(grid_for(R.fine_pts), R.nfine_recs), dim3(kBlock), 0,
                           stream_, (const RegFineRec *)R.fine_recs, (const DevFab *)R.tab_cfp,
                           (const int *)R.C->owner_, R.C->dom_, (const uint8_t *)R.flag, r,
                           (const double *)dt_dev_, fac, R.dx1f);
    }
    CAMR_HIP(hipGetLastError());
    return CAMR_OK;
}

// average_down of the fine level's S_new onto the coarse level's
int average_down(HydroLevel &C, HydroLevel &F, int ratio) {
    int rc = check_refines(F.desc_, C.desc_, ratio);
    if (rc) return rc;
    // EB builds average with EB_average_down (CAMR.cpp:1060), volume-fraction
    // weighted over non-covered children: not built
    if (C.ebflags_ || F.ebflags_) {
        set_error("average_down with embedded boundaries is not built");
        return CAMR_ERR_PARAM;
    }
    for (const Box &B : F.box_)
        for (int d = 0; d < 3; ++d)
            if (B.lo[d] % ratio || (B.hi[d] + 1) % ratio) {
                set_error("fine boxes must be aligned to the refinement ratio");
                return CAMR_ERR_SHAPE;
            }
    if ((rc = C.resolve()) || (rc = F.resolve())) return rc;
    CAMR_HIP(hipStreamSynchronize(F.stream_));
    const DevFab *ctn = C.a_is_new_ ? C.tab_a_ : C.tab_b_;
    const DevFab *ftn = F.a_is_new_ ? F.tab_a_ : F.tab_b_;
    C.launch_all(kb_avgdown, 0, ctn, ftn, (const int *)F.owner_, F.dom_, ratio);
    C.dt_ready_ = false;
    CAMR_HIP(hipGetLastError());
    return CAMR_OK;
}

// One step of level l of a hierarchy with subcycling (Amr::timeStep as
// CAMR drives it): reset the register to the finer level, CAMR_advance on
// level l (CrseAdd / FineAdd), then `ratio` steps of level l+1 --
// recursively, with their own finer levels -- each with its exterior ghost
// data interpolated from level l in time (StateData::InterpAddBox: old /
// new within (t_new - t_old) 1e-3 of either end, else alpha old + beta
// new) and space (FillPatch); then level l's post_timestep (CAMR.cpp:
// 685-699): reflux, avgDown, clean_state.
int amr_advance_level(const std::vector<HydroLevel *> &L, const std::vector<FluxReg *> &R,
                             size_t l, double dt) {
    HydroLevel &C = *L[l];
    const bool has_fine = l + 1 < L.size();
    int rc = C.resolve();
    if (rc) return rc;
    if (has_fine) {
        // reset (CAMR_advance.cpp:64-66)
        FluxReg &Rg = *R[l];
        hipLaunchKernelGGL(k_fill_flat, dim3(grid_for(Rg.pool_n)), dim3(kBlock), 0, C.stream_,
                           Rg.pool, Rg.pool_n, 0.0);
    }
    const double t_old = C.time_;
    C.defer_post_ = has_fine;
    rc = C.advance(dt);
    // a fused step that trips a density floor is redone here, still deferred
    if (!rc) rc = C.resolve();
    if (!rc && C.fused_ && (C.reg_crse_ || C.reg_fine_)) rc = C.reg_add_fused();
    C.defer_post_ = false;
    if (rc || !has_fine) return rc;
    FluxReg &Rg = *R[l];
    HydroLevel &F = *L[l + 1];
    const double t_new = C.time_;
    const double dtf = dt / Rg.ratio;
    const double teps = (t_new - t_old) * 1.e-3;
    for (int s = 0; s < Rg.ratio; ++s) {
        const double t = F.time_;
        int tmode = 2;
        if (t > t_old - teps && t < t_old + teps) tmode = 1;
        else if (t > t_new - teps && t < t_new + teps) tmode = 0;
        const double ta = (t_new - t) / (t_new - t_old), tb = (t - t_old) / (t_new - t_old);
        if ((rc = F.exterior_from(C, Rg.ratio, tmode, ta, tb, true))) return rc;
        if ((rc = amr_advance_level(L, R, l + 1, dtf))) return rc;
    }
    if ((rc = F.resolve())) return rc;
    CAMR_HIP(hipStreamSynchronize(F.stream_));
    // post_timestep of level l (CAMR.cpp:685-699)
    const DevFab *ctn = C.a_is_new_ ? C.tab_a_ : C.tab_b_;
    C.launch_all(kb_reflux, 0, ctn, (const DevFab *)Rg.tab_crse, (const DevFab *)Rg.tab_cfp);
    if ((rc = average_down(C, F, Rg.ratio))) return rc;
    C.clean_all(ctn, true);
    CAMR_HIP(hipGetLastError());
    return CAMR_OK;
}

// One coarse step of an nlev-level hierarchy; R[l] is the register between
// levels l and l+1, and all levels must be at the same time.
int amr_advance(const std::vector<HydroLevel *> &L, const std::vector<FluxReg *> &R, double dt0) {
    if (L.empty() || R.size() + 1 != L.size()) { set_error("need nlev - 1 flux registers"); return CAMR_ERR_PARAM; }
    if (!(dt0 > 0.0)) { set_error("dt must be positive"); return CAMR_ERR_DT; }
    for (size_t l = 0; l + 1 < L.size(); ++l)
        if (!R[l] || R[l]->C != L[l] || R[l]->F != L[l + 1]) {
            set_error("flux register %d does not join levels %d and %d", (int)l, (int)l, (int)l + 1);
            return CAMR_ERR_PARAM;
        }
    for (size_t l = 1; l < L.size(); ++l)
        if (std::fabs(L[l]->time_ - L[0]->time_) > 1e-12 * std::fmax(1.0, std::fabs(L[0]->time_))) {
            set_error("the levels are at different times");
            return CAMR_ERR_PARAM;
        }
    return amr_advance_level(L, R, 0, dt0);
}

int amr_step(HydroLevel &C, HydroLevel &F, FluxReg &R, double dtc) {
    return amr_advance({&C, &F}, {&R}, dtc);
}
} // namespace camr

// ================================================================ plotfile
// AMReX plotfile of a level hierarchy as CAMR writes it
// (CAMR::writePlotFile, Source/Utils/CAMR_io.cpp:407-591, with Amr's
// header lines and VisMF's one-file-per-level data).  AMReX is absent; the
// layout is restated from its published format:
//   dir/Header            plotfile type, variables, geometry, per-level grids
//   dir/Level_<l>/Cell_H  VisMF header: boxes, byte offsets, per-box min/max
//   dir/Level_<l>/Cell_D_00000  per box a FAB header line, then the data,
//                         component after component, x fastest, native
//                         little-endian IEEE doubles
//   dir/job_info          a short description of the run
namespace camr {
namespace {

constexpr double kSmallDen = 1.e-20;  // Derive.cpp local_small_den (EB builds)

// Pointwise derived variables of Source/Utils/Derive.cpp, non-EB forms
// (EB levels guard the density as the EB build does).
double derive_value(const std::string &name, const double *u, bool eb) {
    const double rho = u[URHO];
    if (name == "x_velocity" || name == "y_velocity" || name == "z_velocity") {
        const int c = UMX + (name[0] - 'x');
        return eb ? u[c] / vmax(rho, kSmallDen) : u[c] / rho;  // CAMR_dervelx/y/z
    }
    if (name == "magvel") {  // CAMR_dermagvel
        const double rhoInv = eb ? 1.0 / vmax(rho, kSmallDen) : 1.0 / rho;
        const double d1 = u[UMX] * rhoInv, d2 = u[UMY] * rhoInv, d3 = u[UMZ] * rhoInv;
        return std::sqrt((d1 * d1) + (d2 * d2) + (d3 * d3));
    }
    if (name == "kineng") {  // CAMR_derkineng
        const double sq = u[UMX] * u[UMX] + u[UMY] * u[UMY] + u[UMZ] * u[UMZ];
        return eb ? 0.5 * sq / vmax(rho, kSmallDen) : 0.5 * sq / rho;
    }
    // EOS::REY2P / REY2Gam of the gamma law (EOS.H:135-145)
    if (name == "pressure") {  // CAMR_derpres
        if (eb && rho <= 0.) return 0.;
        const double rhoInv = 1.0 / rho;
        const double e = u[UEINT] * rhoInv;
        return (kGamma - 1.0) * rho * e;
    }
    if (name == "soundspeed" || name == "MachNumber") {  // CAMR_dersoundspeed / dermachnumber
        if (rho <= 0.) return 0.;
        const double rhoInv = 1.0 / rho;
        const double e = u[UEINT] * rhoInv;
        const double pres = (kGamma - 1.0) * rho * e;
        const double c = std::sqrt(kGamma * pres / rho);
        if (name == "soundspeed") return c;
        const double sq = u[UMX] * u[UMX] + u[UMY] * u[UMY] + u[UMZ] * u[UMZ];
        return std::sqrt(sq) / rho / c;
    }
    return 0.0;
}

// derive_lst order of CAMR_setup.cpp:211-300 for the pointwise variables
const char *const kDeriveOrder[] = {"pressure", "kineng", "soundspeed", "MachNumber",
                                    "x_velocity", "y_velocity", "z_velocity", "magvel"};

std::string box_str(const Box &b) {
    std::ostringstream o;
    o << "((" << b.lo[0] << ',' << b.lo[1] << ',' << b.lo[2] << ") (" << b.hi[0] << ','
      << b.hi[1] << ',' << b.hi[2] << ") (0,0,0))";
    return o.str();
}

bool make_dir(const std::string &d) {
    struct stat st;
    if (stat(d.c_str(), &st) == 0) return S_ISDIR(st.st_mode);
    return mkdir(d.c_str(), 0755) == 0;
}

// A level's MultiFab in VisMF's layout: ldir/<name>_D_00000 holds per box
// a FAB header line and the data (component after component, x fastest,
// little-endian IEEE doubles); ldir/<name>_H the boxes, byte offsets and
// per-box min / max.  fill(u, np, out) turns a box's S_new (NVAR
// components of np cells) into its nvars output components.
template <typename Fill>
int write_vismf(const std::string &ldir, const char *name, HydroLevel &L, int nvars, Fill fill) {
    const std::vector<Box> &bx = L.boxes();
    const std::string dname = std::string(name) + "_D_00000";
    std::ofstream D(ldir + "/" + dname, std::ios::binary);
    if (!D) { set_error("cannot write %s/%s", ldir.c_str(), dname.c_str()); return CAMR_ERR_PARAM; }
    std::vector<long long> offs;
    std::vector<std::vector<double>> mn, mx;
    std::vector<double> u, out;
    for (size_t b = 0; b < bx.size(); ++b) {
        const size_t np = bx[b].npts();
        u.resize(NVAR * np);
        int rc = L.download((int)b, u.data());
        if (rc) return rc;
        fill(u, np, out);
        offs.push_back((long long)D.tellp());
        D << "FAB ((8, (64 11 52 0 1 12 0 1023)),(8, (8 7 6 5 4 3 2 1)))" << box_str(bx[b]) << ' '
          << nvars << '\n';
        D.write(reinterpret_cast<const char *>(out.data()), (std::streamsize)(out.size() * sizeof(double)));
        std::vector<double> lo(nvars), hi(nvars);
        for (int v = 0; v < nvars; ++v) {
            double a = out[(size_t)v * np], z = a;
            for (size_t p = 1; p < np; ++p) {
                a = std::min(a, out[(size_t)v * np + p]);
                z = std::max(z, out[(size_t)v * np + p]);
            }
            lo[v] = a;
            hi[v] = z;
        }
        mn.push_back(lo);
        mx.push_back(hi);
    }
    if (!D) { set_error("write failed in %s", ldir.c_str()); return CAMR_ERR_PARAM; }
    std::ofstream V(ldir + "/" + name + "_H");
    V.precision(17);
    V << 1 << '\n' << 1 << '\n' << nvars << '\n' << 0 << '\n';
    V << '(' << bx.size() << " 0\n";
    for (const Box &b : bx) V << box_str(b) << '\n';
    V << ")\n" << bx.size() << '\n';
    for (size_t b = 0; b < bx.size(); ++b) V << "FabOnDisk: " << dname << ' ' << offs[b] << '\n';
    V << '\n' << bx.size() << ',' << nvars << '\n';
    for (auto &r : mn) { for (double v : r) V << v << ','; V << '\n'; }
    V << '\n' << bx.size() << ',' << nvars << '\n';
    for (auto &r : mx) { for (double v : r) V << v << ','; V << '\n'; }
    if (!V) { set_error("write failed in %s", ldir.c_str()); return CAMR_ERR_PARAM; }
    return CAMR_OK;
}

}  // namespace

int write_plotfile(const char *dir, const camr_plot_desc &pd, const std::vector<HydroLevel *> &lev) {
    if (pd.nlev > 1 && !pd.ref_ratio) { set_error("plotfile: ref_ratio needed"); return CAMR_ERR_PARAM; }
    // variables: the state components (CAMR_setup.cpp:159-198) that are plot
    // variables -- all but rho_<species> unless plot_rhoy (CAMR_io.cpp:
    // 100-111) -- then the requested derived variables in derive_lst order
    static const char *const kState[NVAR] = {"density", "xmom", "ymom", "zmom", "rho_E",
                                             "rho_e", "Temp", "rho_F", "rho_A", "rho_P"};
    std::vector<int> comps;
    std::vector<std::string> names;
    for (int n = 0; n < NVAR; ++n) {
        if (n >= UFS && !pd.plot_rhoy) continue;
        comps.push_back(n);
        names.push_back(kState[n]);
    }
    std::vector<std::string> want;
    if (pd.derive_vars) {
        std::istringstream in(pd.derive_vars);
        std::string w;
        while (in >> w) {
            bool ok = false;
            for (const char *d : kDeriveOrder) ok = ok || w == d;
            if (!ok) { set_error("plotfile: derived variable '%s' is not built", w.c_str()); return CAMR_ERR_PARAM; }
            want.push_back(w);
        }
    }
    std::vector<std::string> derived;
    for (const char *d : kDeriveOrder)
        for (const std::string &w : want)
            if (w == d) { derived.push_back(d); names.push_back(d); break; }
    const int nvars = (int)names.size();

    const std::string root(dir);
    if (!make_dir(root)) { set_error("plotfile: cannot create %s", dir); return CAMR_ERR_PARAM; }
    std::ofstream H(root + "/Header");
    if (!H) { set_error("plotfile: cannot write %s/Header", dir); return CAMR_ERR_PARAM; }
    H.precision(17);
    const int f_lev = pd.nlev - 1;
    H << "HyperCLaw-V1.1\n" << nvars << '\n';
    for (const std::string &n : names) H << n << '\n';
    H << 3 << '\n' << lev[0]->time() << '\n' << f_lev << '\n';
    for (int d = 0; d < 3; ++d) H << pd.prob_lo[d] << ' ';
    H << '\n';
    for (int d = 0; d < 3; ++d) H << pd.prob_hi[d] << ' ';
    H << '\n';
    for (int l = 0; l < f_lev; ++l) H << pd.ref_ratio[l] << ' ';
    H << '\n';
    for (int l = 0; l <= f_lev; ++l)
        H << box_str(make_box(lev[l]->desc().hp.domlo, lev[l]->desc().hp.domhi)) << ' ';
    H << '\n';
    for (int l = 0; l <= f_lev; ++l) H << (pd.level_steps ? pd.level_steps[l] : 0) << ' ';
    H << '\n';
    for (int l = 0; l <= f_lev; ++l) {
        for (int d = 0; d < 3; ++d) H << lev[l]->desc().hp.dx[d] << ' ';
        H << '\n';
    }
    H << 0 << '\n' << "0\n";  // Cartesian coordinates; no boundary data

    for (int l = 0; l <= f_lev; ++l) {
        HydroLevel &L = *lev[l];
        const std::vector<Box> &bx = L.boxes();
        const double *dx = L.desc().hp.dx;
        char lname[32];
        snprintf(lname, sizeof lname, "Level_%d", l);
        const std::string ldir = root + "/" + lname;
        if (!make_dir(ldir)) { set_error("plotfile: cannot create %s", ldir.c_str()); return CAMR_ERR_PARAM; }
        H << l << ' ' << bx.size() << ' ' << L.time() << '\n';
        H << (pd.level_steps ? pd.level_steps[l] : 0) << '\n';
        for (const Box &b : bx)
            for (int d = 0; d < 3; ++d)
                H << pd.prob_lo[d] + dx[d] * b.lo[d] << ' ' << pd.prob_lo[d] + dx[d] * (b.hi[d] + 1) << '\n';
        H << lname << "/Cell\n";

        int rc = write_vismf(ldir, "Cell", L, nvars, [&](const std::vector<double> &u, size_t np,
                                                         std::vector<double> &out) {
            out.assign((size_t)nvars * np, 0.0);
            int c = 0;
            for (int n : comps) {
                std::copy(u.begin() + (size_t)n * np, u.begin() + (size_t)(n + 1) * np,
                          out.begin() + (size_t)c * np);
                ++c;
            }
            double cell[NVAR];
            for (const std::string &dn : derived) {
                for (size_t p = 0; p < np; ++p) {
                    for (int n = 0; n < NVAR; ++n) cell[n] = u[(size_t)n * np + p];
                    out[(size_t)c * np + p] = derive_value(dn, cell, L.is_eb());
                }
                ++c;
            }
        });
        if (rc) return rc;
    }
    std::ofstream J(root + "/job_info");
    J << "===============================================================================\n"
      << " CAMR Job Information\n"
      << "===============================================================================\n"
      << "code: camr_amd (MI355X, FP64)\n"
      << "levels: " << pd.nlev << "\n"
      << "time: " << std::setprecision(17) << lev[0]->time() << "\n";
    if (!H) { set_error("plotfile: header write failed"); return CAMR_ERR_PARAM; }
    return CAMR_OK;
}

// A checkpoint of a level hierarchy for restart (CAMR::checkPoint,
// Source/Utils/CAMR_io.cpp:45-84: AmrLevel's state data plus a CPUtime
// file), in the directory layout of AMReX checkpoints: dir/Header with the
// hierarchy (time, ratios, per-level dt and steps, boxes), the new-time
// state of level l in VisMF layout as dir/Level_<l>/SD_0_New_H / _D_00000,
// all NVAR components bit for bit, and dir/CPUtime.
int write_checkpoint(const char *dir, const camr_checkpoint_desc &cd,
                     const std::vector<HydroLevel *> &lev) {
    const std::string root(dir);
    if (!make_dir(root)) { set_error("checkpoint: cannot create %s", dir); return CAMR_ERR_PARAM; }
    std::ofstream H(root + "/Header");
    if (!H) { set_error("checkpoint: cannot write %s/Header", dir); return CAMR_ERR_PARAM; }
    H.precision(17);
    const int f_lev = cd.nlev - 1;
    H << "CheckPointVersion_1.0\n" << 3 << '\n' << lev[0]->time() << '\n' << f_lev << '\n';
    for (int l = 0; l <= f_lev; ++l)
        H << box_str(make_box(lev[l]->desc().hp.domlo, lev[l]->desc().hp.domhi)) << ' ';
    H << '\n';
    for (int l = 0; l < f_lev; ++l) H << cd.ref_ratio[l] << ' ';
    H << '\n';
    for (int l = 0; l <= f_lev; ++l) H << (cd.dt_level ? cd.dt_level[l] : 0.0) << ' ';
    H << '\n';
    for (int l = 0; l <= f_lev; ++l) H << (cd.level_steps ? cd.level_steps[l] : 0) << ' ';
    H << '\n';
    for (int l = 0; l <= f_lev; ++l) {
        HydroLevel &L = *lev[l];
        char lname[32];
        snprintf(lname, sizeof lname, "Level_%d", l);
        const std::string ldir = root + "/" + lname;
        if (!make_dir(ldir)) { set_error("checkpoint: cannot create %s", ldir.c_str()); return CAMR_ERR_PARAM; }
        H << l << ' ' << L.boxes().size() << ' ' << L.time() << '\n';
        for (const Box &b : L.boxes()) H << box_str(b) << '\n';
        H << lname << "/SD_0_New\n";
        int rc = write_vismf(ldir, "SD_0_New", L, NVAR,
                             [](const std::vector<double> &u, size_t, std::vector<double> &out) { out = u; });
        if (rc) return rc;
    }
    std::ofstream T(root + "/CPUtime");
    T << std::setprecision(15) << 0.0;
    if (!H || !T) { set_error("checkpoint: write failed"); return CAMR_ERR_PARAM; }
    return CAMR_OK;
}
}  // namespace camr

// ================================================================ C-ABI
using namespace camr;

struct camr_level {
    HydroLevel impl;
    explicit camr_level(const camr_level_desc &d) : impl(d) {}
};
struct camr_flux_register {
    FluxReg *impl;
};

extern "C" {

void camr_default_params(camr_hydro_params *p) {
    memset(p, 0, sizeof(*p));
    p->ppm_type = 0;
    p->plm_iorder = 4;
    p->use_pslope = 0;
    p->use_flattening = 1;
    p->transverse_reset_density = 1;
    p->allow_negative_energy = 0;
    p->allow_small_energy = 1;
    p->dual_energy_update_E_from_e = 1;
    p->small = 1.e-8;
    p->small_dens = 1.e-7;
    p->small_pres = 1.e-200;
    p->small_temp = 1.e2;
    p->smallu = 1.e-12;
    p->small_num = 1.e-8;
    p->difmag = 0.1;
    p->dual_energy_eta1 = 1.0e0;
    p->dual_energy_eta2 = 1.0e-4;
    for (int d = 0; d < 3; ++d) p->dx[d] = 1.0;
}

void camr_default_level_desc(camr_level_desc *d) {
    memset(d, 0, sizeof(*d));
    camr_default_params(&d->hp);
    d->add_grav_src = 1;
    d->const_grav = 0.0;
    d->cfl = 0.8;
    d->change_max = 1.1;
    d->init_shrink = 1.0;
    d->max_dt = 1.e200;
    d->fixed_dt = -1.0;
    d->stop_time = -1.0;
    d->engine = 0;
    d->batched = 1;
    d->graph = 0;
    d->worker_streams = 0;
}

const char *camr_last_error(void) { return g_err; }
int camr_version(void) { return 100; }

int camr_hydro_ctoprim(const camr_box *qbx, const camr_array4 *u, const camr_array4 *q,
                       const camr_array4 *qaux, const camr_hydro_params *p, void *stream) {
    if (!qbx || !u || !q || !qaux || !p) { set_error("null argument"); return CAMR_ERR_PARAM; }
    const Box B = make_box(qbx->lo, qbx->hi);
    if (!covers(*u, B, NVAR) || !covers(*q, B, QVAR) || !covers(*qaux, B, NQAUX)) {
        set_error("camr_hydro_ctoprim: arrays do not cover qbx");
        return CAMR_ERR_SHAPE;
    }
    hipStream_t s = (hipStream_t)stream;
    int *st = nullptr;
    CAMR_HIP(hipMallocAsync(&st, sizeof(int), s));
    CAMR_HIP(hipMemsetAsync(st, 0, sizeof(int), s));
    launch_on(k_ctoprim, B, s, view(*u), view(*q), view(*qaux), make_const(*p), st);
    int h = 0;
    CAMR_HIP(hipMemcpyAsync(&h, st, sizeof(int), hipMemcpyDeviceToHost, s));
    CAMR_HIP(hipFreeAsync(st, s));
    CAMR_HIP(hipStreamSynchronize(s));
    if (h & (1 << CAMR_ERR_NEG_ENERGY)) {
        set_error("rho E or rho e not positive (Hydro_ctoprim.H:56,68)");
        return CAMR_ERR_NEG_ENERGY;
    }
    return CAMR_OK;
}

int camr_hydro_srctoprim(const camr_box *qbx, const camr_array4 *q, const camr_array4 *qaux,
                         const camr_array4 *src, const camr_array4 *srcq, void *stream) {
    if (!qbx || !q || !qaux || !src || !srcq) { set_error("null argument"); return CAMR_ERR_PARAM; }
    const Box B = make_box(qbx->lo, qbx->hi);
    if (!covers(*q, B, QVAR) || !covers(*qaux, B, NQAUX) || !covers(*src, B, NVAR) ||
        !covers(*srcq, B, QVAR)) {
        set_error("camr_hydro_srctoprim: arrays do not cover qbx");
        return CAMR_ERR_SHAPE;
    }
    launch_on(k_srctoprim, B, (hipStream_t)stream, view(*q), view(*qaux), view(*src), view(*srcq));
    CAMR_HIP(hipGetLastError());
    return CAMR_OK;
}

// Arguments of one hydro_umdrv call, checked; *n_scratch receives the
// scratch doubles the call needs (the dt word not included).
static int umdrv_args(const camr_box *bx, const camr_array4 *uin, const camr_array4 *dsdt,
                      const camr_array4 *q, const camr_array4 *qaux, const camr_array4 *srcq,
                      const camr_array4 flx[3], const camr_array4 area[3],
                      const camr_array4 *vol, double dt, const camr_hydro_params *p,
                      UmdrvArgs *A, size_t *n_scratch) {
    if (!bx || !uin || !dsdt || !q || !qaux || !flx || !p) { set_error("null argument"); return CAMR_ERR_PARAM; }
    int rc = check_params(*p);
    if (rc) return rc;
    if (!(dt > 0.0)) { set_error("dt must be positive"); return CAMR_ERR_DT; }
    const Box B = make_box(bx->lo, bx->hi);
    const Box G = grow(B, CAMR_NGROW);
    if (!B.ok() || !covers(*q, G, QVAR) || !covers(*qaux, G, NQAUX) || !covers(*uin, grow(B, 1), NVAR) ||
        !covers(*dsdt, B, NVAR)) {
        set_error("camr_hydro_umdrv: q/qaux must cover grow(bx,4), uin grow(bx,1), dsdt bx");
        return CAMR_ERR_SHAPE;
    }
    const bool has_src = srcq && srcq->p;
    if (has_src && !covers(*srcq, G, QVAR)) { set_error("srcq must cover grow(bx,4)"); return CAMR_ERR_SHAPE; }
    bool has_area = area && area[0].p && area[1].p && area[2].p;
    for (int d = 0; d < 3; ++d) {
        if (!covers(flx[d], nodes(B, d), NVAR)) { set_error("flx[%d] must cover the faces of bx", d); return CAMR_ERR_SHAPE; }
        if (has_area && !covers(area[d], nodes(B, d), 1)) { set_error("area[%d] must cover the faces of bx", d); return CAMR_ERR_SHAPE; }
    }
    const bool has_vol = vol && vol->p;
    if (has_vol && !covers(*vol, B, 1)) { set_error("vol must cover bx"); return CAMR_ERR_SHAPE; }
    A->bx = B;
    A->qbox = box_of(*q);
    A->hc = make_const(*p);
    A->uin = view(*uin);
    A->dsdt = view(*dsdt);
    A->q = view(*q);
    A->qaux = view(*qaux);
    A->srcq = has_src ? view(*srcq) : view(*q);
    A->has_src = has_src ? 1 : 0;
    for (int d = 0; d < 3; ++d) {
        A->flx[d] = view(flx[d]);
        A->area[d] = has_area ? view(area[d]) : A->flx[d];
    }
    A->has_area = has_area ? 1 : 0;
    A->area_c[0] = p->dx[1] * p->dx[2];
    A->area_c[1] = p->dx[0] * p->dx[2];
    A->area_c[2] = p->dx[0] * p->dx[1];
    A->vol = has_vol ? view(*vol) : A->dsdt;
    A->has_vol = has_vol ? 1 : 0;
    A->vol_c = p->dx[0] * p->dx[1] * p->dx[2];
    *n_scratch = p->do_mol ? umdrv_mol_scratch_doubles(B) : umdrv_staged_scratch_doubles(B);
    return CAMR_OK;
}

// The kernels of one hydro_umdrv call on `scr` (n doubles of scratch, then
// the dt word).
static int umdrv_run(UmdrvArgs &A, double *scr, size_t n, double dt, int do_mol,
                     hipStream_t s) {
    A.dtp = scr + n;
    put_double(scr + n, dt, s);
    const int rc = do_mol ? umdrv_mol(A, scr, s) : umdrv_staged(A, scr, s);
    if (rc) { set_error("kernel launch failed"); return rc; }
    return CAMR_OK;
}

int camr_scratch_bytes(const camr_box *bx, const camr_hydro_params *p, size_t *bytes) {
    if (!bx || !p || !bytes) { set_error("null argument"); return CAMR_ERR_PARAM; }
    const Box B = make_box(bx->lo, bx->hi);
    if (!B.ok()) { set_error("empty box"); return CAMR_ERR_SHAPE; }
    size_t n = p->do_mol ? umdrv_mol_scratch_doubles(B) : umdrv_staged_scratch_doubles(B);
    if (!p->do_mol && canon_fits(B)) n = std::max(n, fused_umdrv_scratch_doubles(B));
    *bytes = (n + 1) * sizeof(double);
    return CAMR_OK;
}

// camr_hydro_umdrv's scratch: a stream-ordered memory pool per device,
// owned by the library and created on the device's first call.  Its release
// threshold is unbounded, so memory freed at the end of a call stays in the
// pool and later calls of the same or smaller size are served without
// going to the driver; concurrent calls on distinct streams each get their
// own block.
static hipMemPool_t umdrv_pool(int dev) {
    static std::mutex mu;
    static std::vector<hipMemPool_t> pools;
    std::lock_guard<std::mutex> g(mu);
    if ((int)pools.size() <= dev) pools.resize(dev + 1, nullptr);
    if (!pools[dev]) {
        hipMemPoolProps props;
        memset(&props, 0, sizeof(props));
        props.allocType = hipMemAllocationTypePinned;
        props.handleTypes = hipMemHandleTypeNone;
        props.location.type = hipMemLocationTypeDevice;
        props.location.id = dev;
        hipMemPool_t pool = nullptr;
        if (hipMemPoolCreate(&pool, &props) != hipSuccess) return nullptr;
        uint64_t keep = UINT64_MAX;
        (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
        pools[dev] = pool;
    }
    return pools[dev];
}

int camr_hydro_umdrv(const camr_box *bx, const camr_array4 *uin, const camr_array4 *dsdt,
                     const camr_array4 *q, const camr_array4 *qaux, const camr_array4 *srcq,
                     const camr_array4 flx[3], const camr_array4 area[3], const camr_array4 *vol,
                     double dt, const camr_hydro_params *p, void *stream) {
    UmdrvArgs A;
    size_t n = 0;
    int rc = umdrv_args(bx, uin, dsdt, q, qaux, srcq, flx, area, vol, dt, p, &A, &n);
    if (rc) return rc;
    hipStream_t s = (hipStream_t)stream;
    // Godunov boxes on a uniform grid with q over exactly grow(bx, 4) run on
    // the fused engine's kernels; others on the staged ones (same results)
    const bool fused = fused_umdrv_applies(A, p->do_mol);
    if (fused) n = fused_umdrv_scratch_doubles(A.bx);
    int dev = 0;
    CAMR_HIP(hipGetDevice(&dev));
    hipMemPool_t pool = umdrv_pool(dev);
    if (!pool) { set_error("scratch pool creation failed on device %d", dev); return CAMR_ERR_ALLOC; }
    double *scr = nullptr;
    CAMR_HIP(hipMallocFromPoolAsync((void **)&scr, (n + 1) * sizeof(double), pool, s));
    if (fused) {
        A.dtp = scr + n;
        put_double(scr + n, dt, s);
        rc = fused_umdrv(A, scr, s);
        if (rc) set_error("kernel launch failed");
    } else {
        rc = umdrv_run(A, scr, n, dt, p->do_mol, s);
    }
    CAMR_HIP(hipFreeAsync(scr, s));
    return rc;
}

// ---------------------------------------- allocation-free hydro_umdrv
// Scratch of camr_hydro_umdrv_noalloc: a file-scope device array, one copy
// per GPU, held from module load to process end.  12 Mi doubles (96 MiB)
// hold the Godunov temporaries of boxes up to 32^3 cells.  In host code
// `scratch` names a host-side stand-in, never a device address: the device
// address of each GPU's copy comes from hipGetSymbolAddress on that GPU.
constexpr size_t kStaticScratchDoubles = size_t(12) << 20;
static __device__ __align__(256) double scratch[kStaticScratchDoubles];

// Device address of this GPU's `scratch`, looked up on the device's first
// call (with the device current) and kept per device.
static double *static_scratch(int dev) {
    static std::mutex mu;
    static std::vector<double *> ptr;
    std::lock_guard<std::mutex> g(mu);
    if ((int)ptr.size() <= dev) ptr.resize(dev + 1, nullptr);
    if (!ptr[dev]) {
        void *a = nullptr;
        const hipError_t e = hipGetSymbolAddress(&a, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !a) {
            set_error("hipGetSymbolAddress(scratch) failed on device %d: %s", dev,
                      hipGetErrorString(e));
            return nullptr;
        }
        ptr[dev] = (double *)a;
    }
    return ptr[dev];
}

int camr_hydro_umdrv_noalloc(const camr_box *bx, const camr_array4 *uin,
                             const camr_array4 *dsdt, const camr_array4 *q,
                             const camr_array4 *qaux, const camr_array4 *srcq,
                             const camr_array4 flx[3], const camr_array4 area[3],
                             const camr_array4 *vol, double dt, const camr_hydro_params *p,
                             void *stream) {
    UmdrvArgs A;
    size_t n = 0;
    int rc = umdrv_args(bx, uin, dsdt, q, qaux, srcq, flx, area, vol, dt, p, &A, &n);
    if (rc) return rc;
    if (n + 1 > kStaticScratchDoubles) {
        set_error("camr_hydro_umdrv_noalloc: the box needs %zu bytes of scratch, more than "
                  "the %zu-byte static maximum", (n + 1) * sizeof(double),
                  kStaticScratchDoubles * sizeof(double));
        return CAMR_ERR_SHAPE;
    }
    int dev = 0;
    CAMR_HIP(hipGetDevice(&dev));
    double *scr = static_scratch(dev);
    if (!scr) return CAMR_ERR_HIP;
    // every scratch element the kernels read is written earlier in the same
    // call, so the previous call's contents need no clearing
    return umdrv_run(A, scr, n, dt, p->do_mol, (hipStream_t)stream);
}

int camr_hydro_umdrv_eb(const camr_box *bx, const camr_array4 *uin, const camr_array4 *dsdt,
                        const camr_array4 *q, const camr_array4 *qaux, const camr_array4 *srcq,
                        const camr_array4 flx[3], const camr_eb_geom *geom, double dt,
                        const camr_hydro_params *p, const int32_t is_periodic[3], void *stream) {
    if (!bx || !uin || !dsdt || !q || !qaux || !flx || !geom || !p || !is_periodic || !geom->flags) {
        set_error("null argument");
        return CAMR_ERR_PARAM;
    }
    int rc = check_params(*p);
    if (rc) return rc;
    if (!(dt > 0.0)) { set_error("dt must be positive"); return CAMR_ERR_DT; }
    if (p->ppm_type != 0 || p->use_pslope != 0) {
        set_error("hydro_umdrv_eb needs ppm_type = 0 and use_pslope = 0 (CAMR.cpp:186-190)");
        return CAMR_ERR_PARAM;
    }
    if (p->do_mol) {
        set_error("hydro_umdrv_eb with the method of lines (MOL_umeth_eb) is not built");
        return CAMR_ERR_PARAM;
    }
    const Box B = make_box(bx->lo, bx->hi);
    const Box G = grow(B, 6), G7 = grow(B, 7);
    if (!B.ok() || !covers(*q, G, QVAR) || !covers(*qaux, G, NQAUX) || !covers(*uin, G, NVAR) ||
        !covers(*dsdt, B, NVAR) || !make_box(geom->flo, geom->fhi).contains(G7) ||
        !covers(geom->vfrac, G7, 1)) {
        set_error("camr_hydro_umdrv_eb: q/qaux/uin must cover grow(bx,6), dsdt bx, the geometry grow(bx,7)");
        return CAMR_ERR_SHAPE;
    }
    const bool has_src = srcq && srcq->p;
    if (has_src && !covers(*srcq, G, QVAR)) { set_error("srcq must cover grow(bx,6)"); return CAMR_ERR_SHAPE; }
    for (int d = 0; d < 3; ++d) {
        if (!covers(flx[d], nodes(B, d), NVAR)) { set_error("flx[%d] must cover the faces of bx", d); return CAMR_ERR_SHAPE; }
        if (!covers(geom->ap[d], nodes(G7, d), 1) || !covers(geom->fc[d], nodes(G7, d), 2)) {
            set_error("ap[%d] / fc[%d] must cover the faces of grow(bx,7)", d, d);
            return CAMR_ERR_SHAPE;
        }
    }
    UmdrvEbArgs A;
    A.bx = B;
    A.qbox = box_of(*q);
    A.hc = make_const(*p);
    A.hc.eb = 1;
    for (int d = 0; d < 3; ++d) A.periodic[d] = is_periodic[d];
    A.uin = view(*uin);
    A.dsdt = view(*dsdt);
    A.q = view(*q);
    A.qaux = view(*qaux);
    A.srcq = has_src ? view(*srcq) : view(*q);
    A.has_src = has_src ? 1 : 0;
    for (int d = 0; d < 3; ++d) A.flx[d] = view(flx[d]);
    const Box fb = make_box(geom->flo, geom->fhi);
    A.g.f.p = geom->flags;
    A.g.f.lo0 = fb.lo[0]; A.g.f.lo1 = fb.lo[1]; A.g.f.lo2 = fb.lo[2];
    A.g.f.js = fb.len(0);
    A.g.f.ks = (int64_t)fb.len(0) * fb.len(1);
    A.g.vf = view(geom->vfrac);
    for (int d = 0; d < 3; ++d) { A.g.ap[d] = view(geom->ap[d]); A.g.fc[d] = view(geom->fc[d]); }
    hipStream_t s = (hipStream_t)stream;
    double *scr = nullptr;
    const size_t n = umdrv_eb_scratch_doubles(B, A.qbox);
    CAMR_HIP(hipMallocAsync((void **)&scr, (n + 1) * sizeof(double), s));
    A.dtp = scr + n;
    put_double(scr + n, dt, s);
    rc = umdrv_eb(A, scr, s);
    CAMR_HIP(hipFreeAsync(scr, s));
    if (rc) { set_error("kernel launch failed"); return rc; }
    return CAMR_OK;
}

int camr_estdt_hydro(const camr_box *bx, const camr_array4 *u, const double dx[3],
                     double *dt_out, void *stream) {
    if (!bx || !u || !dx || !dt_out) { set_error("null argument"); return CAMR_ERR_PARAM; }
    const Box B = make_box(bx->lo, bx->hi);
    if (!covers(*u, B, NVAR)) { set_error("u must cover bx"); return CAMR_ERR_SHAPE; }
    hipStream_t s = (hipStream_t)stream;
    struct { unsigned long long m; int st; int pad; } h = {0x7fefffffffffffffULL, 0, 0};
    char *buf = nullptr;
    CAMR_HIP(hipMallocAsync((void **)&buf, 16, s));
    CAMR_HIP(hipMemcpyAsync(buf, &h, 16, hipMemcpyHostToDevice, s));
    launch_on(k_estdt, B, s, view(*u), dx[0], dx[1], dx[2], (unsigned long long *)buf, (int *)(buf + 8),
              EbFlags{});
    CAMR_HIP(hipMemcpyAsync(&h, buf, 16, hipMemcpyDeviceToHost, s));
    CAMR_HIP(hipFreeAsync(buf, s));
    CAMR_HIP(hipStreamSynchronize(s));
    if (h.st & (1 << CAMR_ERR_DT)) { set_error("ERROR: dt needs to be positive."); return CAMR_ERR_DT; }
    double v;
    memcpy(&v, &h.m, sizeof(v));
    *dt_out = v;
    return CAMR_OK;
}

int camr_level_create(const camr_level_desc *d, camr_level **out) {
    if (!d || !out) { set_error("null argument"); return CAMR_ERR_PARAM; }
    *out = nullptr;
    camr_level *L = new camr_level(*d);
    int rc = L->impl.create();
    if (rc) { delete L; return rc; }
    *out = L;
    return CAMR_OK;
}
int camr_level_destroy(camr_level *L) {
    if (L) { (void)L->impl.sync(); delete L; }
    return CAMR_OK;
}
int camr_level_set_stream(camr_level *L, void *stream) {
    if (!L) return CAMR_ERR_PARAM;
    L->impl.set_stream((hipStream_t)stream);
    return CAMR_OK;
}
int camr_level_upload_box(camr_level *L, int32_t b, const double *h) {
    return L ? L->impl.upload(b, h) : CAMR_ERR_PARAM;
}
int camr_level_download_box(camr_level *L, int32_t b, double *h) {
    return L ? L->impl.download(b, h) : CAMR_ERR_PARAM;
}
int camr_level_state_array(camr_level *L, int32_t b, camr_array4 *out) {
    return (L && out) ? L->impl.state_array(b, out) : CAMR_ERR_PARAM;
}
int camr_level_init_post(camr_level *L) { return L ? L->impl.init_post() : CAMR_ERR_PARAM; }
int camr_level_initial_dt(camr_level *L, double *dt) {
    return (L && dt) ? L->impl.initial_dt(dt) : CAMR_ERR_PARAM;
}
int camr_level_new_dt(camr_level *L, double dt_old, double *dt) {
    return (L && dt) ? L->impl.new_dt(dt_old, dt) : CAMR_ERR_PARAM;
}
int camr_level_advance(camr_level *L, double dt) { return L ? L->impl.advance(dt) : CAMR_ERR_PARAM; }
int camr_level_run(camr_level *L, int32_t nsteps, double *dts) {
    if (!L) return CAMR_ERR_PARAM;
    return L->impl.run(nsteps, dts);
}
int camr_level_sync(camr_level *L) { return L ? L->impl.sync() : CAMR_ERR_PARAM; }
int camr_level_adjoint_forward(camr_level *L, int32_t nsteps, const double *dts,
                               int32_t ckpt_every) {
    return L ? L->impl.adjoint_forward(nsteps, dts, ckpt_every) : CAMR_ERR_PARAM;
}
int camr_level_adj_upload_box(camr_level *L, int32_t b, const double *host) {
    return L ? L->impl.adj_upload(b, host) : CAMR_ERR_PARAM;
}
int camr_level_adj_download_box(camr_level *L, int32_t b, double *host) {
    return L ? L->impl.adj_download(b, host) : CAMR_ERR_PARAM;
}
int camr_level_adj_seed_eb_force(camr_level *L, const double w[3]) {
    if (!L || !w) return CAMR_ERR_PARAM;
    return L->impl.eb_force_seed(w);
}
int camr_level_adjoint_backward(camr_level *L) {
    return L ? L->impl.adjoint_backward() : CAMR_ERR_PARAM;
}
int camr_level_set_exterior(camr_level *L, const double *host) {
    return L ? L->impl.set_exterior(host) : CAMR_ERR_PARAM;
}
int camr_level_exterior_from_coarse(camr_level *fine, camr_level *coarse, int32_t ratio) {
    if (!fine || !coarse || fine == coarse) { set_error("bad level arguments"); return CAMR_ERR_PARAM; }
    return fine->impl.exterior_from(coarse->impl, ratio);
}
int camr_flux_register_create(camr_level *coarse, camr_level *fine, int32_t ratio,
                              camr_flux_register **out) {
    if (!coarse || !fine || coarse == fine || !out) { set_error("bad level arguments"); return CAMR_ERR_PARAM; }
    FluxReg *R = nullptr;
    int rc = flux_register_create(coarse->impl, fine->impl, ratio, &R);
    if (rc) return rc;
    *out = new camr_flux_register{R};
    return CAMR_OK;
}
int camr_flux_register_destroy(camr_flux_register *r) {
    if (!r) return CAMR_ERR_PARAM;
    delete r->impl;
    delete r;
    return CAMR_OK;
}
int camr_level_average_down(camr_level *coarse, camr_level *fine, int32_t ratio) {
    if (!coarse || !fine || coarse == fine) { set_error("bad level arguments"); return CAMR_ERR_PARAM; }
    return average_down(coarse->impl, fine->impl, ratio);
}
int camr_write_checkpoint(const char *dir, const camr_checkpoint_desc *d) {
    if (!d || !dir || d->nlev < 1 || !d->levels) { set_error("checkpoint: bad arguments"); return CAMR_ERR_PARAM; }
    if (d->nlev > 1 && !d->ref_ratio) { set_error("checkpoint: ref_ratio needed"); return CAMR_ERR_PARAM; }
    std::vector<HydroLevel *> lev(d->nlev);
    for (int l = 0; l < d->nlev; ++l) {
        if (!d->levels[l]) { set_error("checkpoint: null level"); return CAMR_ERR_PARAM; }
        lev[l] = &d->levels[l]->impl;
    }
    return write_checkpoint(dir, *d, lev);
}
int camr_level_set_time(camr_level *L, double t) {
    if (!L) { set_error("null argument"); return CAMR_ERR_PARAM; }
    return L->impl.set_time(t);
}
int camr_write_plotfile(const char *dir, const camr_plot_desc *d) {
    if (!d || !dir || d->nlev < 1 || !d->levels) { set_error("plotfile: bad arguments"); return CAMR_ERR_PARAM; }
    std::vector<HydroLevel *> lev(d->nlev);
    for (int l = 0; l < d->nlev; ++l) {
        if (!d->levels[l]) { set_error("plotfile: null level"); return CAMR_ERR_PARAM; }
        lev[l] = &d->levels[l]->impl;
    }
    return write_plotfile(dir, *d, lev);
}
int camr_amr_advance(camr_level *const *levels, camr_flux_register *const *regs, int32_t nlev,
                     double dt0) {
    if (!levels || nlev < 1 || (nlev > 1 && !regs)) { set_error("bad arguments"); return CAMR_ERR_PARAM; }
    std::vector<HydroLevel *> L(nlev);
    std::vector<FluxReg *> R(nlev - 1);
    for (int l = 0; l < nlev; ++l) {
        if (!levels[l]) { set_error("null level"); return CAMR_ERR_PARAM; }
        L[l] = &levels[l]->impl;
    }
    for (int l = 0; l + 1 < nlev; ++l) {
        if (!regs[l] || !regs[l]->impl) { set_error("null flux register"); return CAMR_ERR_PARAM; }
        R[l] = regs[l]->impl;
    }
    return amr_advance(L, R, dt0);
}
int camr_amr_step(camr_level *coarse, camr_level *fine, camr_flux_register *r, double dt_coarse) {
    if (!coarse || !fine || coarse == fine || !r || !r->impl) { set_error("bad arguments"); return CAMR_ERR_PARAM; }
    return amr_step(coarse->impl, fine->impl, *r->impl, dt_coarse);
}
int camr_level_set_eb(camr_level *L, const int32_t *flags, const int32_t flo[3],
                      const int32_t fhi[3]) {
    return L ? L->impl.set_eb(flags, flo, fhi) : CAMR_ERR_PARAM;
}
int camr_level_set_box_eb(camr_level *L, int32_t b, const double *vfrac, const double *apx,
                          const double *apy, const double *apz, const double *fcx,
                          const double *fcy, const double *fcz) {
    if (!L) return CAMR_ERR_PARAM;
    const double *ap[3] = {apx, apy, apz}, *fc[3] = {fcx, fcy, fcz};
    return L->impl.set_box_eb(b, vfrac, ap, fc);
}
int camr_level_set_eb_merge(camr_level *L, int32_t on) {
    return L ? L->impl.set_eb_merge(on) : CAMR_ERR_PARAM;
}
int camr_level_eb_merge_info(camr_level *L, int32_t *nvirtual, int32_t *nmerged) {
    return L ? L->impl.eb_merge_info(nvirtual, nmerged) : CAMR_ERR_PARAM;
}
int camr_level_box_kind(camr_level *L, int32_t b, int32_t *kind) {
    return L ? L->impl.box_kind(b, kind) : CAMR_ERR_PARAM;
}
int camr_level_eb_force(camr_level *L, double force[3]) {
    if (!L || !force) { set_error("null argument"); return CAMR_ERR_PARAM; }
    return L->impl.eb_force(force);
}
int camr_level_set_tuning(camr_level *L, const int32_t block[CAMR_NSTAGE], int32_t order) {
    if (!L || !block) { set_error("null argument"); return CAMR_ERR_PARAM; }
    static_assert(CAMR_NSTAGE == ST_COUNT, "stage kinds");
    Tuning t = default_tuning();
    for (int k = 0; k < ST_COUNT; ++k) t.block[k] = block[k];
    t.order = order;
    return L->impl.set_tuning(t);
}
int camr_level_set_tiles(camr_level *L, const int32_t tile[3 * CAMR_NSTAGE]) {
    if (!L || !tile) { set_error("null argument"); return CAMR_ERR_PARAM; }
    Tuning t;
    int rc = L->impl.get_tuning(&t);
    if (rc) return rc;
    for (int k = 0; k < ST_COUNT; ++k) {
        const int32_t *x = tile + 3 * k;
        if (x[0] == 0 && x[1] == 0 && x[2] == 0) {
            t.tile[k][0] = -1;
            t.tile[k][1] = t.tile[k][2] = 0;
            continue;
        }
        int l[3];
        for (int d = 0; d < 3; ++d) {
            if (x[d] <= 0 || (x[d] & (x[d] - 1))) {
                set_error("stage %d: tile sizes must be powers of two", k);
                return CAMR_ERR_PARAM;
            }
            l[d] = __builtin_ctz((unsigned)x[d]);
            t.tile[k][d] = l[d];
        }
        t.block[k] = x[0] * x[1] * x[2];
        t.zm[k] = 0;  // z-march takes linear blocks
    }
    return L->impl.set_tuning(t);
}
int camr_level_set_xcd(camr_level *L, const int32_t chunk[CAMR_NSTAGE]) {
    if (!L || !chunk) { set_error("null argument"); return CAMR_ERR_PARAM; }
    Tuning t;
    int rc = L->impl.get_tuning(&t);
    if (rc) return rc;
    for (int k = 0; k < ST_COUNT; ++k) t.xcd[k] = chunk[k];
    return L->impl.set_tuning(t);
}
int camr_level_get_xcd(camr_level *L, int32_t chunk[CAMR_NSTAGE]) {
    if (!L || !chunk) { set_error("null argument"); return CAMR_ERR_PARAM; }
    Tuning t;
    int rc = L->impl.get_tuning(&t);
    if (rc) return rc;
    for (int k = 0; k < ST_COUNT; ++k) chunk[k] = t.xcd[k];
    return CAMR_OK;
}
int camr_level_set_zmarch(camr_level *L, const int32_t planes[CAMR_NSTAGE]) {
    if (!L || !planes) { set_error("null argument"); return CAMR_ERR_PARAM; }
    Tuning t;
    int rc = L->impl.get_tuning(&t);
    if (rc) return rc;
    for (int k = 0; k < ST_COUNT; ++k) t.zm[k] = planes[k];
    return L->impl.set_tuning(t);
}
int camr_level_set_ctu(camr_level *L, int32_t on, int32_t planes) {
    if (!L) { set_error("null argument"); return CAMR_ERR_PARAM; }
    Tuning t;
    int rc = L->impl.get_tuning(&t);
    if (rc) return rc;
    t.ctu = on;
    t.ctu_planes = planes;
    return L->impl.set_tuning(t);
}
int camr_level_get_ctu(camr_level *L, int32_t *on, int32_t *planes) {
    if (!L || !on || !planes) { set_error("null argument"); return CAMR_ERR_PARAM; }
    Tuning t;
    int rc = L->impl.get_tuning(&t);
    if (rc) return rc;
    *on = t.ctu;
    *planes = t.ctu_planes;
    return CAMR_OK;
}
int camr_level_get_zmarch(camr_level *L, int32_t planes[CAMR_NSTAGE]) {
    if (!L || !planes) { set_error("null argument"); return CAMR_ERR_PARAM; }
    Tuning t;
    int rc = L->impl.get_tuning(&t);
    if (rc) return rc;
    for (int k = 0; k < ST_COUNT; ++k) planes[k] = t.zm[k];
    return CAMR_OK;
}
int camr_level_get_tiles(camr_level *L, int32_t tile[3 * CAMR_NSTAGE]) {
    if (!L || !tile) { set_error("null argument"); return CAMR_ERR_PARAM; }
    Tuning t;
    int rc = L->impl.get_tuning(&t);
    if (rc) return rc;
    for (int k = 0; k < ST_COUNT; ++k)
        for (int d = 0; d < 3; ++d)
            tile[3 * k + d] = t.tile[k][0] < 0 ? 0 : (1 << t.tile[k][d]);
    return CAMR_OK;
}
int camr_level_get_tuning(camr_level *L, int32_t block[CAMR_NSTAGE], int32_t *order) {
    if (!L || !block || !order) { set_error("null argument"); return CAMR_ERR_PARAM; }
    Tuning t;
    int rc = L->impl.get_tuning(&t);
    if (rc) return rc;
    for (int k = 0; k < ST_COUNT; ++k) block[k] = t.block[k];
    *order = t.order;
    return CAMR_OK;
}
int camr_level_profile(camr_level *L, int32_t enable, const char *kernel_filter) {
    if (!L) return CAMR_ERR_PARAM;
    L->impl.prof().enable(enable != 0, kernel_filter);
    L->impl.prof().reset();
    return CAMR_OK;
}
int camr_level_profile_report(camr_level *L, char *buf, int64_t buflen) {
    if (!L || !buf || buflen <= 0) return CAMR_ERR_PARAM;
    std::string s = L->impl.prof().json();
    if ((int64_t)s.size() + 1 > buflen) { set_error("report buffer too small"); return CAMR_ERR_PARAM; }
    memcpy(buf, s.c_str(), s.size() + 1);
    return CAMR_OK;
}
double camr_level_time(const camr_level *L) { return L ? L->impl.time() : 0.0; }

} // extern "C"
