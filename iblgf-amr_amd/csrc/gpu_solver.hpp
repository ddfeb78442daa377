// Device mirror of the block octree and the driver of the LGF solves.
//
// GpuSolver keeps the reference's solver entry points, on one GPU with the
// tree resident:
//   apply_lgf  PoissonSolver::apply_lgf (include/iblgf/solver/poisson/poisson.hpp:587-894)
//   apply_if   PoissonSolver::apply_lgf_IF / apply_if (poisson.hpp:427-494)
//   fmm_pass   Fmm::apply (include/iblgf/fmm/fmm.hpp:729-870)
// The work of one FMM pass (block lists, sibling groups of the spectral
// multiply-accumulate) is built once from the octree at upload and stays on
// the device.
#pragma once

#include "host/octree.hpp"
#include "kernels.hpp"

#include <cstdint>
#include <functional>
#include <map>
#include <memory>
#include <string>
#include <vector>

namespace iblgf_amd {

template<class T>
struct DevArray
{
    T*     p = nullptr;
    size_t n = 0;
    DevArray() = default;
    DevArray(const DevArray&) = delete;
    DevArray& operator=(const DevArray&) = delete;
    DevArray(DevArray&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
    DevArray& operator=(DevArray&& o) noexcept
    {
        if (this != &o)
        {
            release();
            p = o.p;
            n = o.n;
            o.p = nullptr;
            o.n = 0;
        }
        return *this;
    }
    ~DevArray() { release(); }
    void release();
    void alloc(size_t count);
    void upload(const std::vector<T>& v);
    int  size() const { return static_cast<int>(n); }
};

struct MacLevelPlan
{
    int                lambda = 0;  // level of the targets
    int                d = 0;       // level difference of the kernel
    DevArray<MacGroup> groups;
    DevArray<MacSrc>   srcs;
    DevArray<int>      olist;   // kernel-spectrum rows (offsets) the entries use
    DevArray<uint4>    rowoff;  // per entry: LDS byte offsets of its 8 rows
    DevArray<int>      queue;   // work-queue counters of the launch (mac_policy 2: 8 XCD ranges + done), zero between launches
    int                nused = 0;
    bool               compact = false;  // stage only the olist rows
    long               npairs = 0;
    long               nentries = 0;
    long               nfull = 0;          // entries acting on every target of their group
    long               pairs_full = 0;     // pairs carried by those entries
    long               masked_pop[8] = {};  // other entries by target count 1..8
};

struct PassPlan
{
    int                              L = 0, nl = 0, idx = 0;
    bool                             neighbor_only = false;
    DevArray<int>                    zero_s, zero_t, fwd, inv;
    int                              ninv_top = 0;     // inv starts with the level-L targets
    int                              nfwd_coarse = 0;  // fwd ends with the level-L sources
    std::vector<DevArray<int>>       antrp;  // by level (parents)
    std::vector<DevArray<int>>       intrp;  // by level (parents at level)
    std::vector<MacLevelPlan>        mac;
    long                             npairs = 0;
    long                             ntransfer = 0;  // parent-child transfers (anterp + interp)
    long                             ntransfer_antrp = 0;
};

struct SolveCounts
{
    long passes = 0, n_src = 0, n_tgt = 0, n_pair = 0, n_transfer = 0, n_leaf = 0;
    long n_entry = 0, n_entry_full = 0, n_pair_full = 0, masked_pop[8] = {};
};

// Kernel classes timed with HIP events on the solver's stream.
enum KernelClass
{
    KC_FORWARD = 0,   // forward block DFT
    KC_MAC = 1,       // spectral multiply-accumulate
    KC_INVERSE = 2,   // inverse block DFT
    KC_TRANSFER = 3,  // anterpolation / interpolation / coarsify
    KC_OTHER = 4,     // copies, halo, correction stencil
    KC_COUNT = 5
};

// Event pairs around launches of each kernel class; read after a sync.
class KernelTimer
{
  public:
    ~KernelTimer();
    void enable(bool on);
    bool enabled() const { return on_; }
    void begin(hipStream_t s);
    void end(hipStream_t s, int kind, long work);
    void collect();  // call after the stream is synchronised
    void reset();
    double total_ms[KC_COUNT] = {0, 0, 0, 0, 0};
    long   launches[KC_COUNT] = {0, 0, 0, 0, 0};
    long   work[KC_COUNT] = {0, 0, 0, 0, 0};

  private:
    hipEvent_t next_event();
    bool                    on_ = false;
    std::vector<hipEvent_t> pool_;
    size_t                  used_ = 0;
    struct Rec
    {
        int    kind;
        size_t ev;
        long   work;
    };
    std::vector<Rec> recs_;
};

// The immersed-boundary part of lin_sys_with_ib_solve
// (ifherk.hpp:1451-1517), supplied to the IF-HERK step by the IB system.
struct StepIb
{
    virtual ~StepIb() = default;
    // LinSysSolver::ib_solve (linsys.hpp:113-129): the stage's force from the
    // velocity fields (the IB2xIB integrating factor applied): their
    // projection less the boundary velocity, the fields cleaned, CG_solve
    // from the stage's previous force, kept as that when write_prev.
    virtual void force_solve(const int fields[3], double alpha, int stage, bool write_prev) = 0;
    // smearing (linsys.hpp:350-376) of the current force, added into the
    // fields without cleaning them.
    virtual void smear_add(const int fields[3]) = 0;
};

class GpuSolver
{
  public:
    explicit GpuSolver(int device);
    ~GpuSolver();

    void upload_tree(const Octree& t);
    // Independent trees of one block extent, base level and dx_base in one
    // slot table (level-major, then tree, then key): every solve covers all
    // trees in each launch, each tree's result that of its own solve.
    void upload_forest(const std::vector<const Octree*>& trees);
    int  num_trees() const { return ntrees_; }
    // Combined slot of each octant of tree i (-1 without data).
    const std::vector<int>& forest_slot_of(int i) const { return forest_slot_of_.at(i); }

    int     field_create();
    // An FP32 field ([nslots][n^3] float): accepted by the AMR2AMR solves
    // (apply_lgf / apply_if, batched, sequences, allocation-free), which
    // convert it to and from their FP64 work arrays; every other operation
    // takes FP64 fields only.
    int     field_create_f32();
    bool    field_is_f32(int id) const;
    double* field_ptr(int id);        // FP64 fields only
    void*   field_data(int id);       // either precision
    void    field_upload(int id, const double* host);
    void    field_download(int id, double* host);
    void    field_upload_f32(int id, const float* host);
    void    field_download_f32(int id, float* host);
    size_t  field_size() const { return static_cast<size_t>(nb_) * n3_; }

    // Laplace solve of one scalar component (AMR2AMR, use_correction,
    // subtract_non_leaf = false).  dst is written on leaf blocks only.
    void apply_lgf(int src, int dst);
    // Integrating factor e^{alpha Lap} of one scalar component; dst is
    // written on every block with data.
    void apply_if(int src, int dst, double alpha_base);
    // The same solves of k components (src[c] -> dst[c], e.g. the three
    // vorticity or velocity components) with every launch covering all k;
    // each component's result equals its single-component solve.
    // mesh: kMeshCell (scalar cell fields), kMeshFace or kMeshEdge (component
    // c is field index c % 3): the mesh type the level transfers of the solve
    // (coarsening, interpolation, correction) take their relative positions
    // from, as Source::mesh_type() in the reference (poisson.hpp:632, 1177).
    void apply_lgf_batch(const int* src, const int* dst, int k, int mesh = kMeshCell);
    void apply_if_batch(const int* src, const int* dst, int k, double alpha_base, int mesh = kMeshCell);
    // Size the work arrays for k components ahead of a batched solve.
    void reserve_components(int k);
    // Batched solves issued in order (each as apply_lgf_batch / apply_if_batch);
    // with graphs on, the whole sequence is one captured graph, e.g. every
    // solve of one time step.
    struct SolveItem
    {
        bool             laplace = true;
        double           alpha = 0.0;
        int              mesh = kMeshCell;
        std::vector<int> src, dst;
    };
    void solve_sequence(const std::vector<SolveItem>& items);
    // Times apply_lgf_batch(src, dst, k) (Laplace, or IF with alpha_base)
    // under every combination of the launch shapes that do not change results
    // (multiply-accumulate grid policy, transform grid policy, one or two
    // streams per FMM pass), `reps` eager solves each after a warm-up, keeps
    // the fastest for later solves and returns a report of every timing.
    std::string autotune(const int* src, const int* dst, int k, bool laplace, double alpha_base, int reps);
    void        set_launch_shapes(int mac_policy, int dft_grid);
    int         mac_policy() const { return mac_policy_; }
    int         dft_grid() const { return dft_grid_; }
    bool        concurrency() const { return concurrent_; }
    // A single-component solve that allocates nothing: its four work fields
    // and two spectrum arrays are carved from a caller's device region
    // (region_bytes() bytes, each part 256-byte aligned), which the solve
    // clears by kernel first.  Needs prepared(): every pass plan and kernel
    // spectrum of the solve built beforehand in the current precision.
    size_t region_bytes() const;
    bool   prepared(bool laplace, double alpha_base) const;
    void   solve_in_region(int src, int dst, bool laplace, double alpha_base, void* region, size_t bytes);
    static constexpr int kMaxComponents = 64;
    // Batched solves replayed from captured HIP graphs: the first call with a
    // given (kernel, fields, k, precision, alpha) prepares its plans and
    // kernel spectra, then records its launches into a graph; later calls
    // launch the graph (one launch per solve).  The cache is dropped when the
    // work arrays grow, the tree is re-uploaded or graphs are switched off.
    // Calls run eagerly while kernel timing is on.
    void set_graphs(bool on);
    // Run each FMM pass's coarse chain on a second stream (default on); the
    // results are bit for bit those of one stream.
    void set_concurrency(bool on);
    bool graphs() const { return graphs_; }
    // Build the kernel spectra that a solve will use (cold-path work).
    void prepare_laplace();
    void prepare_if(double alpha_base);

    // Finest-level solves with an immersed-boundary FMM mask (MASK_IB2xIB or
    // MASK_xIB2IB): PoissonSolver::apply_lgf / apply_if with fmm_type set to
    // it (poisson.hpp:587-894, 435-494).  The source is the interior of the
    // finest-level blocks; there is no coarsening and one FMM pass.
    // apply_lgf_masked writes the finest leaves, apply_if_masked every
    // finest-level block (buffers included).
    void apply_lgf_masked(int src, int dst, int mask_type);
    // apply_lgf with fmm_type IB2AMR (LinSysSolver::pressure_correction,
    // linsys.hpp:131-145): finest interiors coarsened, one pass per level with
    // the immersed-boundary blocks as sources, every level's leaves written.
    void apply_lgf_ib2amr(int src, int dst);
    void apply_if_masked(int src, int dst, double alpha_base, int mask_type);
    // The same for k components through every launch (src[c] -> dst[c]; a
    // component may be solved in place, src[c] == dst[c]).
    void apply_if_masked_batch(const int* src, const int* dst, int k, double alpha_base, int mask_type);

    // Block operators over the blocks with data at one level (BlockOp in
    // ops.hpp; scale multiplies 1/dx for the curl transpose).
    void block_op(int op, const int* src, const int* dst, int level, double scale);
    void buffer_exchange(int f, int level);
    // Operator::levelDivergence / levelGradient (operators.hpp:522-538, 588-604)
    void level_divergence(const int src[3], int dst, int level);
    void level_gradient(int src, const int dst[3], int level);
    // Operator::clean_leaf_correction_boundary (operators.hpp:329-513) and
    // clean_ib_region_boundary (73-187)
    void clean_leaf_correction_boundary(int f, int level, bool leaf_only, int width);
    void clean_ib_region_boundary(int f, int level, int width);
    // Operator::domainClean (42-68) and Operator::add (3136-3153)
    void domain_clean(int f);
    void add(int from, int to, double scale);
    // Field operations of the IF-HERK integrator (Ifherk, ifherk.hpp) over
    // every block: to = from * scale (copy, 1944-1960), f = f + v on every
    // node (add_field_expression with a constant, operators.hpp:2985-3002),
    // clean<F>(true) (1306-1350: the one-point ring of the leaves and every
    // non-leaf block zeroed), the ramped boundary clean of
    // Ifherk::clean_leaf_correction_boundary (1352-1412), up_and_down
    // (991-1013: clean, coarsen, interpolate into the correction blocks) of k
    // components of mesh type `mesh` (component c is field index c % 3), and
    // Operator::nonlinear on one level (operators.hpp:2455-2500).
    void copy_scaled(int from, int to, double scale);
    void add_constant(int f, double v);
    void clean_nonleaf(int f);
    void smooth_boundary(int f, int level, bool leaf_only);
    // Operator::maxnorm and blockRootMeanSquare of the k <= 3 component field
    // f[0..k) on every block with data, in slot order (either output may be null).
    void block_norms(const int* f, int k, double* maxnorm, double* rms);
    void up_and_down(const int* f, int k, int mesh);
    void nonlinear(const int u[3], const int w[3], const int d[3], int level, double scale);
    // Ifherk::time_step (ifherk.hpp:1078-1200) without immersed-boundary
    // points, in place on the face velocity u[3]; p (or -1) receives the
    // pressure; frame[9] (or null) the frame velocity of each stage, constant
    // over the domain.  With graphs on, the whole step is one captured graph.
    // refresh 0: the step opens with up_and_down of u; 1: with pad_velocity
    // (1521-1575) and adapt_corr_time_step (1206-1304); 2: pad_velocity only.
    // With ib, the step runs the force solves of lin_sys_with_ib_solve
    // (host-checked CG, so the step runs eagerly, graphs or not).
    void ifherk_step(const int u[3], int p, double dt, double Re, const double* frame, int refresh = 0,
        StepIb* ib = nullptr);
    // Mirror refresh after the host tree was adapted (Octree::adapt): the
    // tree re-uploaded (plans and lists rebuilt, kernel spectra kept), every
    // field id kept; the fields fields[0..nf) (FP64, one field group such as
    // the components of u, field index comp[i], mesh type mesh) carried over
    // block by block by key and the new leaves filled as Ifherk::adapt does
    // (ifherk.hpp:1047-1074): per field, halo fills of the group at every
    // level, then each parent of a new leaf (refinement level > 0)
    // interpolated into all its children, in new-leaf key order.  Other
    // fields restart at zero.  Returns the number of parents interpolated.
    int readapt(const Octree& t, const int* fields, const int* mesh, const int* comp, int nf);
    // Ifherk::up (ifherk.hpp:999-1004) after clean<F>(true): k fields of mesh
    // type mesh coarsened into their parents.
    void coarsen(const int* f, int k, int mesh);
    const Octree& tree() const { return tree_; }
    const std::vector<int>& slot_of() const { return slot_of_; }
    // 0: plain FP64; 1: mixed (source and kernel spectra stored FP32 and
    // multiplied in FP32 with FP64 accumulation; everything else FP64).
    void set_precision(int p);
    int  precision() const { return precision_; }

    void        sync();
    int         device() const { return device_; }
    hipStream_t stream() const { return stream_; }
    int         nblocks() const { return nb_; }
    int         block_n() const { return n_; }
    SolveCounts lgf_counts() const;
    int         num_leaves() const { return n_leaf_; }
    KernelTimer& timer() { return timer_; }

  private:
    void      build_pass(PassPlan& p, int L, int idx, bool neighbor_only);
    PassPlan& pass(int L, int nl, bool neighbor_only);
    PassPlan& pass_idx(int L, int idx, bool neighbor_only);
    static int pass_key(int L, int idx, bool neighbor_only) { return (L * 64 + idx) * 2 + (neighbor_only ? 1 : 0); }
    bool      has_ghat(bool laplace, int d, double alpha) const;
    const DevArray<int>& face_clean_list(int kind, int level, bool leaf_only);
    void      run_pass(PassPlan& p, double scale, bool laplace, double alpha, const Batch& b);
    void      lgf_batch_body(const int* src, const int* dst, int k);
    void      if_batch_body(const int* src, const int* dst, int k, double alpha_base);
    void      run_captured(const std::vector<int64_t>& key, const std::function<void()>& prepare,
             const std::function<void()>& body);
    std::vector<int64_t> batch_key(int kind, const int* src, const int* dst, int k, double alpha) const;
    void      drop_graphs();
    void      check_batch(const int* src, const int* dst, int k);
    Batch     batch(int k) const;
    double2*    ghat_laplace(int d);
    double2*    ghat_if(double alpha);
    const void* ghat(bool laplace, int d, double alpha);  // in the current precision
    void       fill_zero(double* f, const Batch& b = {});
    int        forward_legs() const;  // arithmetic of the forward / inverse transforms' legs
    int        inverse_legs() const;
    double*    F(int id) { return fields_.at(id).p; }
    FieldTable field_table(const int* ids, int k);           // the fields ids[0..k)
    FieldTable work_table(double* base, int k) const;        // components of a work array

    int         device_ = 0;
    hipStream_t stream_ = nullptr;
    hipStream_t side_ = nullptr;  // the coarse chain of an FMM pass
    hipEvent_t  ev_fork_ = nullptr, ev_join_ = nullptr;
    bool        concurrent_ = true;
    int         mac_policy_ = 0, dft_grid_ = 0;  // launch shapes (Batch)
    int         mesh_ = kMeshCell;               // mesh type of the solve in progress
    bool        uploaded_ = false;
    KernelTimer timer_;

    // tree mirror (host copies)
    Octree                  tree_;
    int                     nb_ = 0, n_ = 0, e_ = 0, n3_ = 0, nf_ = 0;
    int                     base_level_ = 0, depth_ = 0, n_leaf_ = 0;
    double                  dx_base_ = 1.0;
    int                     ntrees_ = 1;
    std::vector<int>        slot_of_;   // octant index -> slot (first tree)
    std::vector<std::vector<int>> forest_slot_of_;  // per tree: octant index -> slot
    std::vector<int>        tree_of_;   // slot -> tree
    std::vector<int>        oct_of_;    // slot -> octant index
    std::vector<int>        lvl_begin_;
    std::vector<uint64_t>   key_;
    std::vector<int>        level_, coord_, bbase_, parent_, children_, neighbor_;
    std::vector<uint32_t>   flags_;
    std::vector<uint64_t>   msrc_, mtgt_;
    std::vector<int>        infl_, infl_num_;
    DevArray<int>           d_neighbor_;

    // transfer matrices [2][n*n]
    DevArray<double> m_nli_anterp_, m_nli_interp_, m_cc_restrict_, m_cc_interp_;
    int              intrp_order_ = 5;

    // fields (an FP32 field keeps its floats in a DevArray<double> of half the count)
    std::vector<DevArray<double>> fields_;
    std::vector<char>             field_f32_;
    // Work arrays of the solves, k components each, component-major
    // (component c of a field array at c * field_size(), of a spectrum array
    // at c * nb * NF): source and target temporaries, correction, the FMM
    // pass's source/target and the source/target spectra.
    struct Work
    {
        DevArray<double>  src, tgt, fmm_s, fmm_t;
        DevArray<double2> sspec, tspec;
        int               k = 0;
    };
    Work work_;
    // Where the solves find their work arrays: work_'s arrays, or a caller's
    // scratch region for the duration of an allocation-free solve.  The
    // spectra hold double2 (precision 0) or float2 values.
    struct WorkPtrs
    {
        double* src = nullptr;
        double* tgt = nullptr;
        double* fmm_s = nullptr;
        double* fmm_t = nullptr;
        void*   sspec = nullptr;
        void*   tspec = nullptr;
    };
    WorkPtrs wp_;
    bool                                          graphs_ = false;
    std::map<std::vector<int64_t>, hipGraphExec_t> graph_cache_;

    // kernel spectra
    DevArray<double>                         d_lgf_table_;
    int                                      lgf_nmax_ = 100;
    std::map<int, DevArray<double2>>         ghat_lap_;
    std::map<double, DevArray<double2>>      ghat_if_;
    std::map<int, DevArray<float2>>          ghat_lap32_;
    std::map<double, DevArray<float2>>       ghat_if32_;
    int                                      precision_ = 0;
    DevArray<double>                         gen_sample_;
    DevArray<double2>                        gen_a_, gen_b_;
    DevArray<double>                         d_iv_;

    // plans
    std::map<int, std::unique_ptr<PassPlan>> passes_;
    DevArray<int>                            p_leaves_;
    DevArray<int>                            p_nonleaf_;  // non-leaf blocks at levels >= base
    std::vector<DevArray<int>>               p_coarsify_;    // by level: parent + 8 children
    std::vector<DevArray<int>>               p_level_slots_; // by level
    std::vector<DevArray<int>>               p_leaf_level_;  // leaves by level
    std::vector<DevArray<int>>               p_ccint_;       // children of level-l octants
    std::vector<DevArray<int>>               p_ccint_corr_;  // correction children only
    std::vector<DevArray<int>>               p_kids_;        // child slots of level-l octants
    DevArray<int>                            p_all_;
    std::map<int, DevArray<int>>             clean_lists_;   // {slot, face mask} entries
    std::map<int, DevArray<int>>             corr_level_;    // correction blocks per level

    // IF-HERK step state
    struct StepFields
    {
        int q[3], r[3], g[3], w1[3], w2[3], ui[3], fa[3], ea[3], sf[3], fa2[3], di, ca, ca2;
    };
    StepIb* step_ib_ = nullptr;  // during an IB step
    void                 masked_body(int src, int dst, int mask_type);
    void                 step_prepare(const double alpha[3], int refresh);
    void                 step_body(const int u[3], int p, double dt, const double alpha[3], const double* frame,
                        int refresh);
    void                 step_up(const int* f, int k, int mesh);
    void                 step_down(const int* f, int k, int mesh);
    void                 step_pad_velocity(const int u[3]);
    void                 step_adapt_corr(const int u[3]);
    const DevArray<int>& noncorr_list(int level);
    std::map<int, DevArray<int>> noncorr_lists_;  // non-correction blocks per level
    void                 step_nonlinear(const int src[3], const int dst[3], double scale, const double frame[3]);
    void                 step_lin_sys(double alpha, int stage, bool write_prev);
    const DevArray<int>& corr_list(int level);
    const DevArray<int>& smooth_list(bool leaf_only);
    void                 step_lists();
    bool                 step_ready_ = false;
    StepFields           sf_{};
    DevArray<int>        p_nonleaf_all_;  // non-leaf blocks at every level
    DevArray<int>        p_leaf_ring_;    // {leaf slot, all six sides}
    DevArray<double>     d_ramp_;         // smooth2zero ramp [2][n]
    std::map<int, DevArray<int>> smooth_lists_;  // {base-level slot, 27-bit mask}
};

}  // namespace iblgf_amd
