// engine.hip - context, device layout and the C-ABI of the sweep engine
// (declared in include/ether_gpu.h).
//
// The reference keeps its lattice and spins in module `init`
// (src/init.f90:20-98) and calls Metropolis / overrelaxation / get_tot_energy
// / parallel_tempering once per MCS per MPI rank (src/ether.f90:159-203).
// Here one context holds every replica of every temperature in HBM and each
// of those calls becomes one or a few kernel launches over all replicas.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/ether_gpu.h"
#include "engine_common.h"
#include "philox.h"

namespace ether {
void set_global_error(const std::string& s);
const std::string& global_error();
int sq_scratch_setup(int device, std::string& err);   // sq.hip
}  // namespace ether

using namespace ether;

// longest schedule period (MCS) captured as one graph
constexpr long long kMaxGraphPeriod = 1024;

struct ether_ctx {
    int model = 0, N = 0, R = 0;
    int device = 0;       // the HIP device current at ether_ctx_create
    uint64_t seed = 0;
    std::string err;
    // host-side model (module init)
    int nsp = 0, nshell = 0, cap = 0;
    std::vector<int> species;
    std::vector<double> stagger;
    std::vector<int> nn_count, nn_idx, colour_in;
    std::vector<double> jexc, spin;
    int zee = 0;
    double h[3] = {0, 0, 0}, g = 2.0, mub = 0.0;
    int sia = 0;
    std::vector<double> siav;
    int angle = 0;
    double dphi = 0.0;
    int compat = 0;
    bool lattice_set = false, exchange_set = false, committed = false, betas_set = false;
    // layout
    int ncol = 0;
    std::vector<int> cbeg;
    std::vector<int> slot_of_site, site_of_slot;
    int D = 0;
    int Rp = 0;           // replicas padded to a multiple of 64 (replica-major planes)
    int n_classes = 0;
    std::vector<float> jclass;
    std::vector<double> ising_j;
    std::vector<int> ising_Z, ising_stride;
    int ising_P = 1, ising_offset = 0, ising_table = 0;
    int rec_ncol = 0;     // colour phases to emit while a batch records (its largest count)
    uint64_t graph_gen = 0;  // bumped whenever captured graphs are dropped
    int xcd_map = 2;      // launch parameters of the sweep phases (ether_set_launch)
    int spw = kDefaultSitesPerWave;
    int lps = kWave;      // lanes per site of the sweep (64: one site per wave); set at commit
    int lanes_req = 0;    // ether_set_lanes (0: chosen from R)
    int meas_tiles = 0, meas_sites = 0, meas_rec = 0;
    // Ising sweeps with the lattice resident in LDS (k_ising_resident):
    // possible when a word of every site plus the thresholds fit the budget
    int resident = 1;       // ether_set_resident
    int dft_resident = 1;   // ether_set_dft_resident
    bool res_ok = false;
    int res_block = 256, res_thr_lds = 0;
    size_t res_lds = 0;
    // one workgroup per replica instead of per word (k_ising_resident1):
    // chosen when it takes fewer rounds of workgroups over the compute units
    bool res_rep = false;
    int res1_thr_lds = 0;
    size_t res1_lds = 0;
    // in a batch (k_ising_resident1_batch): one workgroup per replica with
    // its spins and threshold table in up to kBatchResidentLds
    bool res1b_ok = false;
    int res1b_block = 256;
    size_t res1b_lds = 0;
    // spiral-state partners (ether_set_spiral): site-space rows, -1 padded
    int spiral_cap = 0;
    double spiral_count = 0.0;
    std::vector<int> spiral_rows;
    // device
    uint32_t* d_ell = nullptr;
    uint32_t* d_ell_m = nullptr;    // measurement rows (d_ell itself where a site bonds to itself)
    uint32_t* d_ell_m_n = nullptr;  // their lengths
    int D_m = 0, bond_factor = 1;
    // fused measurement (XYZ): possible when every bond has its mirror with
    // the same coupling and no site bonds to itself; meas_lower[q] = 1 where
    // colour q's neighbours are all of earlier colours, 2 where some are
    // (then the phases write per-workgroup records, meas_records); mscale =
    // fixed-point scales (exchange, Zeeman, SIA, moment)
    bool meas_fused_ok = false;
    bool meas_records = false;
    std::vector<int> meas_lower;
    double mscale[4] = {1.0, 1.0, 1.0, 1.0};
    unsigned long long* d_macc = nullptr;
    long long* d_mrec = nullptr;    // [ncol][G][kWave][mrec_pp][kMeasWords]
    int mrec_cap = 0;               // units per phase allocated
    int mrec_pp = 0;                // unit slots per (phase, group, lane): ceil(mrec_cap / G)
    uint8_t* d_sp = nullptr;
    float* d_w = nullptr;
    float* d_jclass = nullptr;
    float* d_zvec = nullptr;
    float* d_dsia = nullptr;
    int* d_stride = nullptr;
    uint32_t* d_thr = nullptr;
    double* d_ising_j = nullptr;
    double* d_ising_zc = nullptr;
    float* d_sx = nullptr;
    float* d_sy = nullptr;
    float* d_sz = nullptr;
    float* d_phi = nullptr;
    int8_t* d_s = nullptr;
    int* d_slot_of = nullptr;
    int* d_site_of_slot = nullptr;   // site of each slot (the structure factor's line form)
    int* d_rep_at = nullptr;
    float* d_beta_f = nullptr;
    double* d_beta = nullptr;
    unsigned long long* d_acc = nullptr;
    double* d_mom = nullptr;
    double* d_partial = nullptr;
    double* d_Ept = nullptr;
    double* d_parts = nullptr;
    double* d_M = nullptr;
    long long* d_bonds = nullptr;
    int* d_spiral = nullptr;
    double* d_mom_sp = nullptr;
    double* d_psi = nullptr;
    int n_groups = 0;
    int* d_gstart = nullptr;
    unsigned long long* d_pt_acc = nullptr;
    unsigned long long* d_pt_att = nullptr;
    unsigned long long* d_trips = nullptr;
    int* d_trip_state = nullptr;
    int* d_nacc = nullptr;
    float* d_scratch_h = nullptr;
    float* d_scratch_e = nullptr;
    std::vector<double> beta_h;
    uint32_t step = 0, init_count = 0, pt_count = 0;
    hipStream_t stream = nullptr;
    // a batch's completion that work on this context must wait for, made a
    // wait on the stream at the context's next use (cst): made at once, it
    // would put a barrier on a hardware queue that another batch's stream may
    // share and hold that batch back
    hipEvent_t after = nullptr;
    // device counters d_ctr = {sweep base, swap base, swap count} and their host
    // mirrors; kernels use counter + baked offset so a captured graph replays
    uint32_t* d_ctr = nullptr;
    uint32_t dev_step = 0, dev_pt = 0, dev_sc = 0;
    bool capturing = false;
    uint32_t cap_step0 = 0, cap_pt0 = 0;
    int use_graphs = 1;
    struct GraphEntry {
        int phase, P, coa, ovrr_iv, pt_iv, n_pt;
        double ovrr_p;
        hipGraph_t graph;
        hipGraphExec_t exec;
    };
    std::vector<GraphEntry> graphs;
    // kernel timing: event pairs per kind, reused from a pool
    bool timing = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev_pool;
    size_t ev_used = 0;
    std::vector<int> ev_kind;
    std::vector<int> ev_count;
    std::vector<void*> allocs;
    char* d_arena = nullptr;   // the block holding every buffer of the committed context
    size_t arena_bytes = 0;
    // MLP / S(q) state lives in the extension files
    void* mlp = nullptr;
    void* sq = nullptr;
};

// The context's stream, after the wait for a batch that ran the context
// (ether_batch_run leaves the wait pending until the context is next used).
static inline hipStream_t cst(ether_ctx* c) {
    if (c->after) {
        (void)hipStreamWaitEvent(c->stream, c->after, 0);
        c->after = nullptr;
    }
    return c->stream;
}

// --------------------------------------------------------------------------
// helpers
// --------------------------------------------------------------------------
static int fail(ether_ctx* c, int code, const std::string& msg) {
    if (c) c->err = msg;
    else set_global_error(msg);
    return code;
}

#define HIPCHK(ctx, expr)                                                                   \
    do {                                                                                    \
        hipError_t _e = (expr);                                                             \
        if (_e != hipSuccess)                                                               \
            return fail(ctx, ETHER_ERR_HIP,                                                 \
                        std::string(#expr) + " failed: " + hipGetErrorString(_e));          \
    } while (0)

template <class T>
static int dev_alloc(ether_ctx* c, T** p, size_t n) {
    if (n == 0) n = 1;
    hipError_t e = hipMalloc((void**)p, n * sizeof(T));
    if (e != hipSuccess)
        return fail(c, ETHER_ERR_HIP, std::string("hipMalloc failed: ") + hipGetErrorString(e));
    c->allocs.push_back(*p);
    return ETHER_OK;
}

// Host to device, complete on return: the runtime may read a pageable source
// after hipMemcpyAsync returns (when the copy reaches the head of the stream),
// and the callers pass buffers that are freed as soon as they return.
template <class T>
static int upload(ether_ctx* c, T* d, const T* h, size_t n) {
    HIPCHK(c, hipMemcpyAsync(d, h, n * sizeof(T), hipMemcpyHostToDevice, cst(c)));
    HIPCHK(c, hipStreamSynchronize(cst(c)));
    return ETHER_OK;
}

// One device allocation carved into 256-byte aligned buffers, laid out in
// three runs: tables (copied from one host image), buffers that start at zero
// (one fill) and buffers the kernels write before they read.  Pointers are
// assigned when commit() has allocated the block; the block is freed with the
// context.
class Arena {
  public:
    template <class T>
    void table(T** p, const std::vector<T>& h) {
        place((void**)p, h.size() * sizeof(T));
        src_.push_back({off_.back(), h.data(), h.size() * sizeof(T)});
    }
    template <class T>
    void zeroed(T** p, size_t n) { place((void**)p, n * sizeof(T)); }
    template <class T>
    void plain(T** p, size_t n) { place((void**)p, n * sizeof(T)); }
    void end_tables() { tables_end_ = top_; }
    void end_zeroed() { zero_end_ = top_; }
    int commit(ether_ctx* c) {
        char* base = nullptr;
        hipError_t e = hipMalloc((void**)&base, std::max<size_t>(top_, 256));
        if (e != hipSuccess)
            return fail(c, ETHER_ERR_HIP, std::string("hipMalloc failed: ") + hipGetErrorString(e));
        c->allocs.push_back(base);
        c->d_arena = base;
        c->arena_bytes = std::max<size_t>(top_, 256);
        for (size_t k = 0; k < dst_.size(); ++k) *dst_[k] = base + off_[k];
        if (tables_end_) {
            image_.assign(tables_end_, 0);
            for (const Src& s : src_) std::memcpy(image_.data() + s.off, s.h, s.bytes);
            // the image is freed with the Arena: wait for the copy
            HIPCHK(c, hipMemcpyAsync(base, image_.data(), tables_end_, hipMemcpyHostToDevice, cst(c)));
            HIPCHK(c, hipStreamSynchronize(cst(c)));
        }
        if (zero_end_ > tables_end_)
            HIPCHK(c, hipMemsetAsync(base + tables_end_, 0, zero_end_ - tables_end_, cst(c)));
        return ETHER_OK;
    }
  private:
    struct Src {
        size_t off;
        const void* h;
        size_t bytes;
    };
    void place(void** p, size_t bytes) {
        dst_.push_back(p);
        off_.push_back(top_);
        top_ += (std::max<size_t>(bytes, 1) + 255) / 256 * 256;
    }
    std::vector<void**> dst_;
    std::vector<size_t> off_;
    std::vector<Src> src_;
    std::vector<char> image_;
    size_t top_ = 0, tables_end_ = 0, zero_end_ = 0;
};

// The acceptance counters are kAccCopies copies of [R] (the sweeps spread
// their atomics over them): an asynchronous copy of all of them, and the
// total of one slot once the stream has been synchronised.
static int copy_acc_async(ether_ctx* c, std::vector<unsigned long long>& raw) {
    raw.resize((size_t)kAccCopies * c->R);
    HIPCHK(c, hipMemcpyAsync(raw.data(), c->d_acc, sizeof(unsigned long long) * raw.size(),
                             hipMemcpyDeviceToHost, cst(c)));
    return ETHER_OK;
}
static unsigned long long acc_total(const std::vector<unsigned long long>& raw, int R, int r) {
    unsigned long long t = 0;
    for (int k = 0; k < kAccCopies; ++k) t += raw[(size_t)k * R + r];
    return t;
}

#define TRY(expr)                      \
    do {                               \
        int _rc = (expr);              \
        if (_rc != ETHER_OK) return _rc; \
    } while (0)

static int need_committed(ether_ctx* c) {
    if (!c) return fail(nullptr, ETHER_ERR_ARG, "null context");
    if (!c->committed) return fail(c, ETHER_ERR_STATE, "context not committed (call ether_commit)");
    return ETHER_OK;
}
static int need_betas(ether_ctx* c) {
    TRY(need_committed(c));
    if (!c->betas_set) return fail(c, ETHER_ERR_STATE, "temperatures not set (call ether_set_betas)");
    return ETHER_OK;
}

// --------------------------------------------------------------------------
// life cycle
// --------------------------------------------------------------------------
extern "C" int ether_ctx_create(int32_t model, int32_t n_sites, int32_t n_replicas, uint64_t seed,
                                ether_ctx** out) {
    if (!out) return fail(nullptr, ETHER_ERR_ARG, "ether_ctx_create: null output pointer");
    *out = nullptr;
    if (model != ETHER_MODEL_ISING && model != ETHER_MODEL_XYZ)
        return fail(nullptr, ETHER_ERR_ARG, "ether_ctx_create: model must be ising (0) or xyz (1)");
    if (n_sites <= 0 || n_sites > (1 << 24) || n_replicas <= 0)
        return fail(nullptr, ETHER_ERR_ARG, "ether_ctx_create: bad n_sites or n_replicas");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
        return fail(nullptr, ETHER_ERR_NODEV, "ether_ctx_create: no HIP device visible");
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) != hipSuccess)
        return fail(nullptr, ETHER_ERR_NODEV, "ether_ctx_create: hipGetDeviceProperties failed");
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(nullptr, ETHER_ERR_NODEV,
                    std::string("ether_ctx_create: device is ") + prop.gcnArchName +
                        ", this library is built for gfx950 only");
    {
        // the device array of the allocation-free structure factor: its
        // address on this device, looked up once
        std::string err;
        const int rc = sq_scratch_setup(dev, err);
        if (rc != ETHER_OK) return fail(nullptr, rc, "ether_ctx_create: " + err);
    }
    ether_ctx* c = new ether_ctx();
    c->device = dev;
    c->model = model;
    c->N = n_sites;
    c->R = n_replicas;
    c->seed = seed;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return fail(nullptr, ETHER_ERR_HIP, "ether_ctx_create: hipStreamCreate failed");
    }
    *out = c;
    return ETHER_OK;
}

extern void ether_ext_destroy(ether_ctx* c);  // mlp / sq state

extern "C" void ether_ctx_destroy(ether_ctx* c) {
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(cst(c));   // and for a batch that ran the context
    ether_ext_destroy(c);
    for (auto& g : c->graphs) {
        (void)hipGraphExecDestroy(g.exec);
        (void)hipGraphDestroy(g.graph);
    }
    for (auto& e : c->ev_pool) {
        (void)hipEventDestroy(e.first);
        (void)hipEventDestroy(e.second);
    }
    for (void* p : c->allocs) (void)hipFree(p);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

extern "C" int ether_last_error(const ether_ctx* c, char* buf, int32_t len) {
    if (!buf || len <= 0) return ETHER_ERR_ARG;
    const std::string& s = c ? c->err : global_error();
    std::snprintf(buf, (size_t)len, "%s", s.c_str());
    return ETHER_OK;
}

// --------------------------------------------------------------------------
// model set-up
// --------------------------------------------------------------------------
extern "C" int ether_set_lattice(ether_ctx* c, int32_t n_species, const int32_t* species,
                                 const double* stagger, int32_t n_shells, int32_t capacity,
                                 const int32_t* nn_count, const int32_t* nn_idx,
                                 const int32_t* colour) {
    if (!c) return fail(nullptr, ETHER_ERR_ARG, "null context");
    if (c->committed) return fail(c, ETHER_ERR_STATE, "ether_set_lattice after commit");
    if (n_species <= 0 || n_species > kMaxSpecies)
        return fail(c, ETHER_ERR_LIMIT, "ether_set_lattice: n_species must be 1..16");
    if (n_shells <= 0 || capacity <= 0 || !species || !stagger || !nn_count || !nn_idx)
        return fail(c, ETHER_ERR_ARG, "ether_set_lattice: bad arguments");
    const int N = c->N;
    c->nsp = n_species;
    c->nshell = n_shells;
    c->cap = capacity;
    c->species.assign(species, species + N);
    c->stagger.assign(stagger, stagger + N);
    c->nn_count.assign(nn_count, nn_count + (size_t)n_shells * N);
    c->nn_idx.assign(nn_idx, nn_idx + (size_t)n_shells * N * capacity);
    for (int i = 0; i < N; ++i)
        if (c->species[i] < 1 || c->species[i] > n_species)
            return fail(c, ETHER_ERR_ARG, "ether_set_lattice: species id out of range");
    for (size_t q = 0; q < c->nn_count.size(); ++q) {
        if (c->nn_count[q] < 0 || c->nn_count[q] > capacity)
            return fail(c, ETHER_ERR_ARG, "ether_set_lattice: neighbour count out of range");
        const int32_t* nb = nn_idx + q * capacity;
        for (int k = 0; k < c->nn_count[q]; ++k)
            if (nb[k] < 0 || nb[k] >= N)
                return fail(c, ETHER_ERR_ARG, "ether_set_lattice: neighbour index out of range");
    }
    if (colour) c->colour_in.assign(colour, colour + N);
    else c->colour_in.clear();
    c->lattice_set = true;
    return ETHER_OK;
}

extern "C" int ether_set_exchange(ether_ctx* c, const double* j_exc, const double* spin) {
    if (!c) return fail(nullptr, ETHER_ERR_ARG, "null context");
    if (!c->lattice_set) return fail(c, ETHER_ERR_STATE, "ether_set_exchange before ether_set_lattice");
    if (c->committed) return fail(c, ETHER_ERR_STATE, "ether_set_exchange after commit");
    if (!j_exc || !spin) return fail(c, ETHER_ERR_ARG, "ether_set_exchange: null pointer");
    c->jexc.assign(j_exc, j_exc + (size_t)c->nshell * c->nsp * c->nsp * 3);
    c->spin.assign(spin, spin + c->nsp);
    c->exchange_set = true;
    return ETHER_OK;
}

extern "C" int ether_set_zeeman(ether_ctx* c, int32_t on, const double* h3, double g_factor,
                                double mu_b) {
    if (!c) return fail(nullptr, ETHER_ERR_ARG, "null context");
    if (c->committed) return fail(c, ETHER_ERR_STATE, "ether_set_zeeman after commit");
    c->zee = on ? 1 : 0;
    if (on) {
        if (!h3) return fail(c, ETHER_ERR_ARG, "ether_set_zeeman: null field");
        for (int x = 0; x < 3; ++x) c->h[x] = h3[x];
        c->g = g_factor;
        c->mub = mu_b;
    }
    return ETHER_OK;
}

extern "C" int ether_set_sia(ether_ctx* c, int32_t on, const double* sia) {
    if (!c) return fail(nullptr, ETHER_ERR_ARG, "null context");
    if (c->committed) return fail(c, ETHER_ERR_STATE, "ether_set_sia after commit");
    if (!c->lattice_set) return fail(c, ETHER_ERR_STATE, "ether_set_sia before ether_set_lattice");
    // SIA enters only for XYZ (src/Hamiltonian.f90:50, src/get_sia_values.f90:32)
    c->sia = (on && c->model == ETHER_MODEL_XYZ) ? 1 : 0;
    if (c->sia) {
        if (!sia) return fail(c, ETHER_ERR_ARG, "ether_set_sia: null vector");
        c->siav.assign(sia, sia + 3 * c->nsp);
    }
    return ETHER_OK;
}

extern "C" int ether_set_proposal(ether_ctx* c, int32_t angle, double dphi_rad) {
    if (!c) return fail(nullptr, ETHER_ERR_ARG, "null context");
    if (c->committed) return fail(c, ETHER_ERR_STATE, "ether_set_proposal after commit");
    c->angle = (angle && c->model == ETHER_MODEL_XYZ) ? 1 : 0;
    c->dphi = dphi_rad;
    return ETHER_OK;
}

static void drop_graphs(ether_ctx* c);

extern "C" int ether_set_compat(ether_ctx* c, int32_t flags) {
    if (!c) return fail(nullptr, ETHER_ERR_ARG, "null context");
    if (c->compat != flags) drop_graphs(c);
    c->compat = flags;
    return ETHER_OK;
}

// --------------------------------------------------------------------------
// commit: classes, colouring, slots, ELL, device buffers
// --------------------------------------------------------------------------
namespace {

struct Edge {
    int j;
    int cls;
};

// Greedy colouring of the interaction graph visiting vertices in `order`.
int greedy_colour(int N, const std::vector<int>& adj_ptr, const std::vector<int>& adj,
                  const std::vector<int>& order, std::vector<int>& col) {
    col.assign(N, -1);
    std::vector<int> mark(N + 1, -1);
    int ncol = 0;
    for (int v : order) {
        for (int q = adj_ptr[v]; q < adj_ptr[v + 1]; ++q) {
            int c = col[adj[q]];
            if (c >= 0) mark[c] = v;
        }
        int c = 0;
        while (mark[c] == v) ++c;
        col[v] = c;
        ncol = std::max(ncol, c + 1);
    }
    return ncol;
}

}  // namespace

// Lanes per site of the sweep: the smallest of 16, 32 (XYZ) or 8, 16, 32
// (Ising, 4 replicas per lane) that holds every replica, 64 otherwise.
// Groups of exactly R lanes (three sites per wave for 17-21 replicas, up to
// eight for 8) were slower on the batched uneven lattices: simple-cubic
// Heisenberg 0.785 against 0.718 ms per MCS, pyrochlore 0.342 against 0.306;
// more sites per wave cost more than the idle lanes they fill.
static int auto_lanes(int model, int R) {
    const int need = model == ETHER_MODEL_ISING ? (R + 3) / 4 : R;
    for (int l = model == ETHER_MODEL_ISING ? 8 : 16; l < kWave; l *= 2)
        if (need <= l) return l;
    return kWave;
}
static int lanes_shift_of(int lanes) {
    int sft = 0;
    while ((1 << sft) < lanes) ++sft;
    return sft;
}

extern "C" int ether_set_lanes(ether_ctx* c, int32_t lanes_per_site) {
    if (!c) return fail(nullptr, ETHER_ERR_ARG, "null context");
    const bool ising = c->model == ETHER_MODEL_ISING;
    const bool pow2 = lanes_per_site > 0 && (lanes_per_site & (lanes_per_site - 1)) == 0;
    if (lanes_per_site != 0 && (lanes_per_site < 1 || lanes_per_site > kWave || (ising && (!pow2 || lanes_per_site < 8))))
        return fail(c, ETHER_ERR_ARG, "ether_set_lanes: lanes per site must be 0 (automatic), 1 to 64 "
                                      "(XYZ) or a power of two from 8 to 64 (Ising)");
    if (lanes_per_site != 0 && lanes_per_site < kWave) {
        const int per = c->model == ETHER_MODEL_ISING ? 4 : 1;
        if (lanes_per_site * per < c->R)
            return fail(c, ETHER_ERR_ARG, "ether_set_lanes: fewer lanes per site than replicas");
    }
    c->lanes_req = lanes_per_site;
    if (c->committed) {
        const int lps = lanes_per_site ? lanes_per_site : auto_lanes(c->model, c->R);
        if (lps != c->lps) drop_graphs(c);
        c->lps = lps;
    }
    return ETHER_OK;
}

extern "C" int ether_commit(ether_ctx* c) {
    if (!c) return fail(nullptr, ETHER_ERR_ARG, "null context");
    if (c->committed) return fail(c, ETHER_ERR_STATE, "context already committed");
    if (!c->lattice_set || !c->exchange_set)
        return fail(c, ETHER_ERR_STATE, "ether_commit needs ether_set_lattice and ether_set_exchange");
    const int N = c->N, nsp = c->nsp, ns = c->nshell, cap = c->cap;
    const bool ising = c->model == ETHER_MODEL_ISING;

    // ---- coupling classes and per-site bonds (zero couplings dropped), CSR ----
    // A bond's coupling is a function of (shell, species of i, species of j):
    // each such triple is resolved to a class once, when the walk over (site,
    // shell, neighbour) first meets it, so classes are numbered in the order
    // their couplings first appear.
    std::map<std::tuple<float, float, float>, int> fcls;
    std::map<double, int> icls;
    c->jclass.assign(3, 0.f);
    c->ising_j.assign(1, 0.0);
    std::vector<int> tcls((size_t)ns * nsp * nsp, -2);   // -2 not met yet, -1 zero coupling
    auto class_of = [&](int s, int a, int b) -> int {
        int& t = tcls[((size_t)s * nsp + a) * nsp + b];
        if (t != -2) return t;
        const double* J = c->jexc.data() + (((size_t)s * nsp + a) * nsp + b) * 3;
        if (ising) {
            if (J[2] == 0.0) return t = -1;
            auto it = icls.find(J[2]);
            if (it != icls.end()) return t = it->second;
            const int cls = (int)c->ising_j.size();
            icls[J[2]] = cls;
            c->ising_j.push_back(J[2]);
            c->jclass.push_back(0.f);
            c->jclass.push_back(0.f);
            c->jclass.push_back((float)J[2]);
            return t = cls;
        }
        if (J[0] == 0.0 && J[1] == 0.0 && J[2] == 0.0) return t = -1;
        auto key = std::make_tuple((float)J[0], (float)J[1], (float)J[2]);
        auto it = fcls.find(key);
        if (it != fcls.end()) return t = it->second;
        const int cls = (int)(c->jclass.size() / 3);
        fcls[key] = cls;
        c->jclass.push_back((float)J[0]);
        c->jclass.push_back((float)J[1]);
        c->jclass.push_back((float)J[2]);
        return t = cls;
    };
    std::vector<int> e_ptr(N + 1, 0);
    std::vector<Edge> edge;
    {
        size_t total = 0;
        for (size_t q = 0; q < (size_t)ns * N; ++q) total += (size_t)c->nn_count[q];
        edge.reserve(total);
    }
    for (int i = 0; i < N; ++i) {
        const int a = c->species[i] - 1;
        for (int s = 0; s < ns; ++s) {
            const int cnt = c->nn_count[(size_t)s * N + i];
            const int32_t* nb = c->nn_idx.data() + ((size_t)s * N + i) * cap;
            for (int k = 0; k < cnt; ++k) {
                const int j = nb[k];
                const int cls = class_of(s, a, c->species[j] - 1);
                if (cls >= 0) edge.push_back({j, cls});
            }
        }
        e_ptr[i + 1] = (int)edge.size();
    }
    c->n_classes = (int)(c->jclass.size() / 3) - 1;
    if (c->n_classes > kMaxClasses)
        return fail(c, ETHER_ERR_LIMIT, "more than 255 distinct exchange couplings");
    if (ising && c->n_classes > kMaxIsingClasses)
        return fail(c, ETHER_ERR_LIMIT, "Ising path supports at most 8 distinct J_zz values");
    if (ising && nsp > kMaxIsingClasses)
        return fail(c, ETHER_ERR_LIMIT, "Ising path supports at most 8 species");

    // ---- colouring ----
    std::vector<int> col;
    if (!c->colour_in.empty()) {
        col = c->colour_in;
        int mx = 0;
        for (int i = 0; i < N; ++i) {
            if (col[i] < 0) return fail(c, ETHER_ERR_ARG, "negative colour");
            mx = std::max(mx, col[i]);
        }
        for (int i = 0; i < N; ++i)
            for (int q = e_ptr[i]; q < e_ptr[i + 1]; ++q)
                if (edge[q].j != i && col[edge[q].j] == col[i])
                    return fail(c, ETHER_ERR_ARG, "colouring joins two interacting sites");
        c->ncol = mx + 1;
    } else {
        // the undirected interaction graph (self bonds dropped), rows sorted
        std::vector<int> adj_ptr(N + 1, 0), adj;
        {
            std::vector<int> deg(N + 1, 0);
            for (int i = 0; i < N; ++i)
                for (int q = e_ptr[i]; q < e_ptr[i + 1]; ++q)
                    if (edge[q].j != i) {
                        deg[i + 1]++;
                        deg[edge[q].j + 1]++;
                    }
            for (int i = 0; i < N; ++i) deg[i + 1] += deg[i];
            std::vector<int> raw(deg[N]), fill(deg.begin(), deg.end() - 1);
            for (int i = 0; i < N; ++i)
                for (int q = e_ptr[i]; q < e_ptr[i + 1]; ++q)
                    if (edge[q].j != i) {
                        raw[fill[i]++] = edge[q].j;
                        raw[fill[edge[q].j]++] = i;
                    }
            adj.reserve(raw.size());
            for (int i = 0; i < N; ++i) {
                std::sort(raw.begin() + deg[i], raw.begin() + deg[i + 1]);
                for (int q = deg[i]; q < deg[i + 1]; ++q)
                    if (q == deg[i] || raw[q] != raw[q - 1]) adj.push_back(raw[q]);
                adj_ptr[i + 1] = (int)adj.size();
            }
        }
        // natural order, then basis-major orders for every plausible basis size
        std::vector<int> order(N), best;
        for (int i = 0; i < N; ++i) order[i] = i;
        int bestn = greedy_colour(N, adj_ptr, adj, order, best);
        for (int k = 2; k <= 64 && bestn > 2; ++k) {
            if (N % k) continue;
            int q = 0;
            for (int l = 0; l < k; ++l)
                for (int cell = 0; cell < N / k; ++cell) order[q++] = cell * k + l;
            std::vector<int> cc;
            int n = greedy_colour(N, adj_ptr, adj, order, cc);
            if (n < bestn) {
                bestn = n;
                best.swap(cc);
            }
        }
        col.swap(best);
        c->ncol = bestn;
    }
    // ---- slots: colour-major, site order within a colour ----
    c->cbeg.assign(c->ncol + 1, 0);
    for (int i = 0; i < N; ++i) c->cbeg[col[i] + 1]++;
    for (int q = 0; q < c->ncol; ++q) c->cbeg[q + 1] += c->cbeg[q];
    c->slot_of_site.assign(N, 0);
    c->site_of_slot.assign(N, 0);
    {
        std::vector<int> fill(c->cbeg.begin(), c->cbeg.end() - 1);
        for (int i = 0; i < N; ++i) {
            int s = fill[col[i]]++;
            c->slot_of_site[i] = s;
            c->site_of_slot[s] = i;
        }
    }
    // ---- ELL ----
    int D = 0;
    for (int i = 0; i < N; ++i) D = std::max(D, e_ptr[i + 1] - e_ptr[i]);
    // rows padded (class 0, zero coupling) to an unrolled width bucket
    {
        const int b = ell_bucket(std::max(D, 1));
        D = b ? b : D;
    }
    c->D = D;
    // slot-major rows: a wave's whole neighbour list is one scalar-loadable run
    std::vector<uint32_t> ell((size_t)std::max(D, 1) * N, 0u);
    for (int i = 0; i < N; ++i) {
        uint32_t* row = ell.data() + (size_t)c->slot_of_site[i] * D;
        for (int q = e_ptr[i]; q < e_ptr[i + 1]; ++q)
            row[q - e_ptr[i]] = (uint32_t)c->slot_of_site[edge[q].j] | ((uint32_t)edge[q].cls << 24);
    }
    // ---- measurement rows: each bond from its endpoint with the smaller site
    // index (in the supercell order each site then holds about half of its
    // bonds), counted twice; a self bond (a lattice one cell thick along a
    // periodic axis) keeps the full rows, counted once.  ell_m_n[slot] is the
    // row's length. ----
    // Booking a bond once and counting it twice is the reference's double sum
    // only if every bond i -> j has its mirror j -> i with the same coupling
    // (src/j_values.f90:83-84 makes j_exc symmetric; neighbour lists built by
    // distance are); otherwise the full rows are used too.
    bool self_bond = false;
    for (int i = 0; i < N && !self_bond; ++i)
        for (int q = e_ptr[i]; q < e_ptr[i + 1]; ++q)
            if (edge[q].j == i) {
                self_bond = true;
                break;
            }
    bool mirrored = !self_bond;
    if (mirrored) {
        // the multiset of bonds (i, j, class) equals that of (j, i, class):
        // row i of the bonds, as (j, class) sorted, equals row i of the
        // transpose, as (i', class) sorted
        std::vector<int> t_ptr(N + 1, 0);
        for (const Edge& ed : edge) t_ptr[ed.j + 1]++;
        for (int i = 0; i < N && mirrored; ++i) {
            if (t_ptr[i + 1] != e_ptr[i + 1] - e_ptr[i]) mirrored = false;
            t_ptr[i + 1] += t_ptr[i];
        }
        if (mirrored) {
            std::vector<uint64_t> tr(edge.size());
            std::vector<int> fill(t_ptr.begin(), t_ptr.end() - 1);
            for (int i = 0; i < N; ++i)
                for (int q = e_ptr[i]; q < e_ptr[i + 1]; ++q)
                    tr[fill[edge[q].j]++] = ((uint64_t)i << 32) | (uint32_t)edge[q].cls;
            std::vector<uint64_t> row;
            for (int i = 0; i < N && mirrored; ++i) {
                row.clear();
                for (int q = e_ptr[i]; q < e_ptr[i + 1]; ++q)
                    row.push_back(((uint64_t)edge[q].j << 32) | (uint32_t)edge[q].cls);
                std::sort(row.begin(), row.end());
                std::sort(tr.begin() + t_ptr[i], tr.begin() + t_ptr[i + 1]);
                mirrored = std::equal(row.begin(), row.end(), tr.begin() + t_ptr[i]);
            }
        }
    }
    std::vector<uint32_t> ell_m, ell_m_n(N, 0u);
    if (!mirrored) {
        c->D_m = D;
        c->bond_factor = 1;
        for (int i = 0; i < N; ++i) ell_m_n[c->slot_of_site[i]] = (uint32_t)(e_ptr[i + 1] - e_ptr[i]);
    } else {
        int Dm = 0;
        for (int i = 0; i < N; ++i) {
            int u = 0;
            for (int q = e_ptr[i]; q < e_ptr[i + 1]; ++q) u += edge[q].j > i;
            ell_m_n[c->slot_of_site[i]] = (uint32_t)u;
            Dm = std::max(Dm, u);
        }
        c->D_m = Dm;
        c->bond_factor = 2;
        ell_m.assign((size_t)std::max(Dm, 1) * N, 0u);
        for (int i = 0; i < N; ++i) {
            uint32_t* row = ell_m.data() + (size_t)c->slot_of_site[i] * Dm;
            int k = 0;
            for (int q = e_ptr[i]; q < e_ptr[i + 1]; ++q)
                if (edge[q].j > i) row[k++] = (uint32_t)c->slot_of_site[edge[q].j] | ((uint32_t)edge[q].cls << 24);
        }
    }
    // ---- Ising index tables ----
    if (ising) {
        const int K = c->n_classes;
        c->ising_Z.assign(K + 1, 0);
        std::vector<int> cnt(K + 1, 0);
        for (int i = 0; i < N; ++i) {
            std::fill(cnt.begin(), cnt.end(), 0);
            for (int q = e_ptr[i]; q < e_ptr[i + 1]; ++q) cnt[edge[q].cls]++;
            for (int k = 1; k <= K; ++k) c->ising_Z[k] = std::max(c->ising_Z[k], cnt[k]);
        }
        c->ising_stride.assign(K + 1, 0);
        long long P = 1;
        c->ising_offset = 0;
        for (int k = 1; k <= K; ++k) {
            c->ising_stride[k] = (int)P;
            c->ising_offset += c->ising_Z[k] * (int)P;
            P *= (2 * c->ising_Z[k] + 1);
            if (P * 2 * nsp > 65536)
                return fail(c, ETHER_ERR_LIMIT, "Ising threshold table exceeds 65536 entries");
        }
        c->ising_P = (int)P;
        c->ising_table = (int)(P * 2 * nsp);
        // the sweep counts a site's down neighbours per class in one byte per
        // replica (SWAR), the measurement its differing bonds per row likewise
        for (int k = 1; k <= K; ++k)
            if (c->ising_Z[k] > 255)
                return fail(c, ETHER_ERR_LIMIT, "Ising path supports at most 255 neighbours of one "
                                                "coupling class per site");
    }
    // ---- per-slot species / magnetisation weights ----
    std::vector<uint8_t> sp(N);
    std::vector<float> w(N);
    for (int i = 0; i < N; ++i) {
        const int s = c->slot_of_site[i];
        sp[s] = (uint8_t)(c->species[i] - 1);
        w[s] = (float)(c->stagger[i] * c->spin[c->species[i] - 1]);
    }
    std::vector<float> zvec(3 * nsp, 0.f), dsia(3 * nsp, 0.f);
    std::vector<double> zc(nsp, 0.0);
    for (int a = 0; a < nsp; ++a) {
        // -g*muB*s*H  (src/Hamiltonian.f90:120)
        double f = c->zee ? -c->g * c->mub * c->spin[a] : 0.0;
        for (int x = 0; x < 3; ++x) zvec[3 * a + x] = (float)(f * c->h[x]);
        zc[a] = f * c->h[2];
        if (c->sia)
            for (int x = 0; x < 3; ++x) dsia[3 * a + x] = (float)c->siav[3 * a + x];
    }
    // ---- fused measurement (XYZ without spiral partners): a phase books the
    // bonds of its sites to neighbours of earlier colours (final already),
    // each bond once from its later-coloured end and counted twice; with two
    // colours a phase has either all its neighbours earlier or none, with
    // three or more some phases have both and book through the field from
    // their earlier neighbours alone, into per-workgroup records.  Fixed-point scales: powers of two with
    // |term| * scale <= 2^48 / N for every site term, so that the sum over
    // all sites stays below 2^48 (kernels round each term to an integer,
    // exactly, below 2^51) ----
    c->meas_fused_ok = !ising && mirrored && !c->spiral_cap;
    c->meas_lower.assign(c->ncol, 0);
    for (int q = 0; q < c->ncol && c->meas_fused_ok; ++q) {
        bool lo = false, hi = false;
        for (int s = c->cbeg[q]; s < c->cbeg[q + 1]; ++s) {
            const int i = c->site_of_slot[s];
            for (int k = e_ptr[i]; k < e_ptr[i + 1]; ++k)
                (c->slot_of_site[edge[k].j] < c->cbeg[q] ? lo : hi) = true;
        }
        c->meas_lower[q] = lo ? (hi ? 2 : 1) : 0;
        if (lo && hi) c->meas_records = true;
    }
    c->meas_records = c->meas_records && c->meas_fused_ok;
    if (c->meas_fused_ok) {
        double bound[4] = {0.0, 0.0, 0.0, 0.0};
        for (int i = 0; i < N; ++i) {
            double t = 0.0;          // |h| <= sum of the largest |J| component per bond
            for (int q = e_ptr[i]; q < e_ptr[i + 1]; ++q) {
                const int k = edge[q].cls;
                t += std::max(std::fabs(c->jclass[3 * k]),
                              std::max(std::fabs(c->jclass[3 * k + 1]), std::fabs(c->jclass[3 * k + 2])));
            }
            bound[0] = std::max(bound[0], 2.0 * t);
        }
        for (int a = 0; a < nsp; ++a) {
            bound[1] = std::max(bound[1], (double)std::fabs(zvec[3 * a]) + std::fabs(zvec[3 * a + 1]) +
                                              std::fabs(zvec[3 * a + 2]));
            bound[2] = std::max(bound[2], (double)std::fabs(dsia[3 * a]) + std::fabs(dsia[3 * a + 1]) +
                                              std::fabs(dsia[3 * a + 2]));
        }
        for (int s = 0; s < N; ++s) bound[3] = std::max(bound[3], (double)std::fabs(w[s]));
        for (int k = 0; k < 4; ++k) {
            // 2x for spins a few ulp longer than 1 and the FP32 roundings of a term
            const double total = 2.0 * (double)N * bound[k];
            c->mscale[k] = total > 0.0 ? std::ldexp(1.0, std::max(-1000, std::min(1000,
                                             (int)std::floor(std::log2(std::ldexp(1.0, 49) / total)))))
                                       : 1.0;
        }
    }
    // ---- device buffers ----
    const int R = c->R;
    c->Rp = (R + kWave - 1) / kWave * kWave;
    const size_t RN = (size_t)c->Rp * N;
    // the sweep kernels address a spin plane with 32-bit buffer offsets
    if (RN * sizeof(float) > (size_t)INT32_MAX)
        return fail(c, ETHER_ERR_LIMIT, "ether_commit: sites x padded replicas exceeds 2^29 "
                                      "(one FP32 spin plane must stay under 2 GiB)");
    const int G = c->Rp / kWave;
    c->lps = c->lanes_req ? c->lanes_req : auto_lanes(c->model, R);
    // measurement: wave units of at least 8 sites each, a workgroup's four
    // waves sharing one tile and one partial record.  Short measurement rows
    // (3 bonds per site on average: simple cubic, triangular) run best with
    // about 16384 units per pass, long ones (pyrochlore's 9) with about 4096,
    // which walk longer runs of consecutive sites (measured at 32^3 SC, 96^2
    // triangular and 16^3 pyrochlore cells: 32 vs 47 us, 18 vs 20 us, 181 vs
    // 278 us).
    long long row_entries = 0;
    for (int s = 0; s < N; ++s) row_entries += ell_m_n[s];
    const long long units = row_entries > 4LL * N ? 4096 : 16384;
    // Ising with 128 or more replicas: a lane holds four replicas, so a group
    // is 256 replicas and there are fewer wave units per site; waves walk
    // from 2 sites up (16^3 cells, 190 replicas: 11.2 us per pass against
    // 18.8 us with one replica per lane).  Fewer replicas keep one per lane
    // and walks from 8 sites, as XYZ.
    const bool wide = ising && c->Rp >= 2 * kWave;     // the launcher's choice of Ising kernel
    const int groups = wide ? (c->Rp + 4 * kWave - 1) / (4 * kWave) : G;
    const int min_sites = wide ? 2 : 8;
    c->meas_sites = kWavesPerBlock * std::max(min_sites, (int)(((long long)N * groups + units - 1) / units));
    if (wide) {
        // a wave adds its sites' per-byte counts of differing bonds into 16-bit
        // fields: at most 65535 / D_m sites per wave, and at most 255 entries
        // in one measurement row
        if (c->D_m > 255)
            return fail(c, ETHER_ERR_LIMIT, "Ising path supports at most 255 neighbours per site");
        const int max_per_wave = std::max(1, 65535 / std::max(c->D_m, 1));
        c->meas_sites = std::min(c->meas_sites, kWavesPerBlock * max_per_wave);
    }
    c->meas_tiles = (N + c->meas_sites - 1) / c->meas_sites;
    c->meas_rec = ising ? (c->n_classes + nsp + 1) : (c->spiral_cap ? 9 : 6);
    // resident Ising sweeps: one word (4 replicas) of every site in LDS, and
    // the four replicas' threshold tables when they fit too
    c->res_ok = false;
    c->res1b_ok = false;
    int maxcol = 0;
    for (int q = 0; q < c->ncol; ++q) maxcol = std::max(maxcol, c->cbeg[q + 1] - c->cbeg[q]);
    // about two sites of the largest colour per thread, 64 to 1024 threads
    const int res_threads = std::max(64, std::min(1024, (maxcol / 2 + 63) / 64 * 64));
    if (ising && c->ncol <= kMaxResColours) {
        const size_t n16 = ((size_t)N + 15) / 16 * 16;
        c->res1b_lds = n16 + (size_t)4 * c->ising_table;
        c->res1b_ok = c->res1b_lds <= kBatchResidentLds;
        c->res1b_block = res_threads;
    }
    if (ising && c->ncol <= kMaxResColours && (size_t)N * 4 <= kResidentLds) {
        c->res_ok = true;
        c->res_thr_lds = ((size_t)N * 4 + (size_t)16 * c->ising_table <= kResidentLds) ? 1 : 0;
        c->res_lds = (size_t)N * 4 + (c->res_thr_lds ? (size_t)16 * c->ising_table : 0);
        c->res_block = res_threads;
        // per replica: N int8 spins (+ one threshold table); a workgroup does
        // about half the instructions of a word's (the Philox block is most of
        // a site's work either way), so it wins while the replicas take fewer
        // than twice the rounds of workgroups the words take
        const size_t n16 = ((size_t)N + 15) / 16 * 16;
        c->res1_thr_lds = (n16 + (size_t)4 * c->ising_table <= kResidentLds) ? 1 : 0;
        c->res1_lds = n16 + (c->res1_thr_lds ? (size_t)4 * c->ising_table : 0);
        int ncu = 0;
        HIPCHK(c, hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, c->device));
        ncu = std::max(ncu, 1);
        const long long rounds1 = ((long long)R + ncu - 1) / ncu, rounds4 = ((long long)c->Rp / 4 + ncu - 1) / ncu;
        c->res_rep = rounds1 < 2 * rounds4;
    }
    // four zero entries past the last row: a 4-wide chunk of the last row stays in the table
    ell.resize(ell.size() + 4, 0u);
    if (!ell_m.empty()) ell_m.resize(ell_m.size() + 4, 0u);
    std::vector<int> ident(R);
    for (int r = 0; r < R; ++r) ident[r] = r;
    std::vector<int> spiral_rows;
    if (c->spiral_cap) {
        // partner rows in slot space, in the order the sites' rows were given
        spiral_rows.assign((size_t)N * c->spiral_cap, -1);
        for (int i = 0; i < N; ++i) {
            const int s = c->slot_of_site[i];
            for (int k = 0; k < c->spiral_cap; ++k) {
                const int j = c->spiral_rows[(size_t)i * c->spiral_cap + k];
                if (j < 0) break;
                spiral_rows[(size_t)s * c->spiral_cap + k] = c->slot_of_site[j];
            }
        }
    }
    // Every buffer of the context in one allocation: the tables built here
    // first (one host image, one copy), then the counters that start at zero
    // (one fill), then the rest.
    Arena ar;
    ar.table(&c->d_ell, ell);
    if (!ell_m.empty()) ar.table(&c->d_ell_m, ell_m);
    ar.table(&c->d_ell_m_n, ell_m_n);
    ar.table(&c->d_sp, sp);
    ar.table(&c->d_site_of_slot, c->site_of_slot);
    ar.table(&c->d_w, w);
    ar.table(&c->d_jclass, c->jclass);
    ar.table(&c->d_zvec, zvec);
    ar.table(&c->d_dsia, dsia);
    ar.table(&c->d_ising_j, c->ising_j);
    ar.table(&c->d_ising_zc, zc);
    ar.table(&c->d_slot_of, ident);
    ar.table(&c->d_rep_at, ident);
    if (ising) ar.table(&c->d_stride, c->ising_stride);
    if (c->spiral_cap) ar.table(&c->d_spiral, spiral_rows);
    ar.end_tables();
    if (ising) ar.zeroed(&c->d_bonds, (size_t)R * (c->n_classes + 1));
    ar.zeroed(&c->d_acc, (size_t)kAccCopies * R);
    ar.zeroed(&c->d_mom, (size_t)R * ETHER_MOM_COUNT);
    ar.zeroed(&c->d_pt_acc, (size_t)R);
    ar.zeroed(&c->d_pt_att, (size_t)R);
    ar.zeroed(&c->d_trips, (size_t)R);
    ar.zeroed(&c->d_trip_state, (size_t)R);
    if (c->meas_fused_ok && !c->meas_records) ar.zeroed(&c->d_macc, (size_t)kMeasCopies * c->Rp * kMeasWords);
    if (c->meas_records) {
        // records per phase at the measuring sweep's shape (at least
        // kMeasRecordSpw site steps per wave, so at least 4 * kMeasRecordSpw
        // sites per workgroup), XCD padding included; zero where a phase
        // writes none (an empty colour)
        int maxcol = 0;
        for (int q = 0; q < c->ncol; ++q) maxcol = std::max(maxcol, c->cbeg[q + 1] - c->cbeg[q]);
        const int per_wg = kWavesPerBlock * kMeasRecordSpw;
        c->mrec_cap = G * ((maxcol + per_wg - 1) / per_wg) + 7;
        c->mrec_pp = (c->mrec_cap + G - 1) / G;
        ar.zeroed(&c->d_mrec, (size_t)c->ncol * G * kWave * c->mrec_pp * kMeasWords);
    }
    if (c->spiral_cap) {
        ar.zeroed(&c->d_mom_sp, (size_t)R * 3);
        ar.zeroed(&c->d_psi, (size_t)R);
    }
    ar.end_zeroed();
    if (ising) {
        ar.plain(&c->d_s, RN);
        ar.plain(&c->d_thr, (size_t)R * c->ising_table);
    } else {
        ar.plain(&c->d_sx, RN);
        ar.plain(&c->d_sy, RN);
        ar.plain(&c->d_sz, RN);
        if (c->angle) ar.plain(&c->d_phi, RN);
    }
    ar.plain(&c->d_beta_f, (size_t)R);
    ar.plain(&c->d_beta, (size_t)R);
    ar.plain(&c->d_partial, (size_t)c->Rp * c->meas_tiles * c->meas_rec);
    ar.plain(&c->d_Ept, (size_t)R);
    ar.plain(&c->d_parts, (size_t)R * 4);
    ar.plain(&c->d_M, (size_t)R * 3);
    ar.plain(&c->d_nacc, 1);
    ar.plain(&c->d_gstart, (size_t)R + 1);
    ar.plain(&c->d_scratch_h, (size_t)N * 4);
    ar.plain(&c->d_scratch_e, (size_t)N);
    ar.plain(&c->d_ctr, 4);
    TRY(ar.commit(c));
    if (ising) {
        // ordered start (all +1), padding replicas included
        HIPCHK(c, hipMemsetAsync(c->d_s, 1, RN, cst(c)));
    } else {
        // every spin (and padding lane) along +z
        HIPCHK(c, hipMemsetAsync(c->d_sx, 0, RN * 4, cst(c)));
        HIPCHK(c, hipMemsetAsync(c->d_sy, 0, RN * 4, cst(c)));
        HIPCHK(c, hipMemsetD32Async(c->d_sz, 0x3f800000, RN, cst(c)));
        if (c->angle) HIPCHK(c, hipMemsetAsync(c->d_phi, 0, RN * 4, cst(c)));
    }
    launch_set_ctr(c->d_ctr, 0, 0, 0, cst(c));
    c->dev_step = c->dev_pt = c->dev_sc = 0;
    HIPCHK(c, hipStreamSynchronize(cst(c)));
    c->committed = true;
    return ETHER_OK;
}

extern "C" int ether_get_layout(ether_ctx* c, int32_t* n_colours, int32_t* colour, int32_t* rng_id) {
    TRY(need_committed(c));
    if (n_colours) *n_colours = c->ncol;
    for (int i = 0; i < c->N; ++i) {
        int s = c->slot_of_site[i];
        if (rng_id) rng_id[i] = s;
        if (colour) {
            int q = 0;
            while (s >= c->cbeg[q + 1]) ++q;
            colour[i] = q;
        }
    }
    return ETHER_OK;
}

// --------------------------------------------------------------------------
// temperatures
// --------------------------------------------------------------------------
static int upload_ising_thresholds(ether_ctx* c) {
    const int K = c->n_classes, nsp = c->nsp, P = c->ising_P;
    std::vector<uint32_t> thr((size_t)c->R * c->ising_table);
    std::vector<int> n(K + 1, 0);
    for (int t = 0; t < c->R; ++t) {
        const double beta = c->beta_h[t];
        for (int a = 0; a < nsp; ++a)
            for (int up = 0; up < 2; ++up) {
                const int sigma = up ? 1 : -1;
                for (int idx = 0; idx < P; ++idx) {
                    // decode n_k from the mixed-radix index
                    int rem = idx;
                    double hfield = 0.0;
                    for (int k = 1; k <= K; ++k) {
                        int digit = rem % (2 * c->ising_Z[k] + 1);
                        rem /= (2 * c->ising_Z[k] + 1);
                        n[k] = digit - c->ising_Z[k];
                        hfield += c->ising_j[k] * n[k];
                    }
                    // dE = (S'-S)_z (h + zeeman), S'_z - S_z = -2 sigma
                    double dE = -2.0 * sigma * hfield;
                    if (c->zee) dE += -c->g * c->mub * c->spin[a] * (-2.0 * sigma) * c->h[2];
                    double p = std::exp(-beta * dE);
                    uint32_t th;
                    if (!(p < 1.0)) th = 0xffffffffu;
                    else {
                        double f = std::floor(p * 4294967296.0);
                        if (f > 4294967294.0) f = 4294967294.0;
                        th = (uint32_t)f;
                    }
                    thr[(size_t)t * c->ising_table + (size_t)(a * 2 + up) * P + idx] = th;
                }
            }
    }
    TRY(upload(c, c->d_thr, thr.data(), thr.size()));
    return ETHER_OK;
}

static int upload_betas(ether_ctx* c, const double* beta) {
    if (!beta) return fail(c, ETHER_ERR_ARG, "ether_set_betas: null pointer");
    std::vector<float> bf(c->R);
    for (int r = 0; r < c->R; ++r) {
        if (!(beta[r] > 0.0) || !std::isfinite(beta[r]))
            return fail(c, ETHER_ERR_ARG, "ether_set_betas: beta must be positive and finite");
        bf[r] = (float)beta[r];
    }
    c->beta_h.assign(beta, beta + c->R);
    TRY(upload(c, c->d_beta, beta, (size_t)c->R));
    TRY(upload(c, c->d_beta_f, bf.data(), (size_t)c->R));
    if (c->model == ETHER_MODEL_ISING) TRY(upload_ising_thresholds(c));
    return ETHER_OK;
}

extern "C" int ether_retune_betas(ether_ctx* c, const double* beta) {
    TRY(need_betas(c));
    TRY(upload_betas(c, beta));
    HIPCHK(c, hipStreamSynchronize(cst(c)));
    return ETHER_OK;
}

extern "C" int ether_set_betas(ether_ctx* c, const double* beta) {
    TRY(need_committed(c));
    TRY(upload_betas(c, beta));
    std::vector<int> ident(c->R);
    for (int r = 0; r < c->R; ++r) ident[r] = r;
    TRY(upload(c, c->d_slot_of, ident.data(), (size_t)c->R));
    TRY(upload(c, c->d_rep_at, ident.data(), (size_t)c->R));
    HIPCHK(c, hipStreamSynchronize(cst(c)));
    c->betas_set = true;
    return ETHER_OK;
}

extern "C" int ether_set_ladders(ether_ctx* c, int32_t n_groups, const int32_t* group_start) {
    TRY(need_committed(c));
    if (n_groups < 0) return fail(c, ETHER_ERR_ARG, "ether_set_ladders: negative group count");
    if (n_groups > 0) {
        if (!group_start) return fail(c, ETHER_ERR_ARG, "ether_set_ladders: null group_start");
        if (group_start[0] != 0 && group_start[0] < 0)
            return fail(c, ETHER_ERR_ARG, "ether_set_ladders: bad group_start");
        for (int g = 0; g < n_groups; ++g)
            if (group_start[g + 1] <= group_start[g] || group_start[g + 1] > c->R)
                return fail(c, ETHER_ERR_ARG, "ether_set_ladders: groups must be increasing and <= R");
        TRY(upload(c, c->d_gstart, group_start, (size_t)n_groups + 1));
    }
    drop_graphs(c);   // the ladder count is baked into captured swap nodes
    c->n_groups = n_groups;
    HIPCHK(c, hipMemsetAsync(c->d_pt_acc, 0, sizeof(unsigned long long) * c->R, cst(c)));
    HIPCHK(c, hipMemsetAsync(c->d_pt_att, 0, sizeof(unsigned long long) * c->R, cst(c)));
    HIPCHK(c, hipMemsetAsync(c->d_trips, 0, sizeof(unsigned long long) * c->R, cst(c)));
    HIPCHK(c, hipMemsetAsync(c->d_trip_state, 0, sizeof(int) * c->R, cst(c)));
    HIPCHK(c, hipStreamSynchronize(cst(c)));
    return ETHER_OK;
}

// --------------------------------------------------------------------------
// argument blocks
// --------------------------------------------------------------------------
// workgroups of a colour phase: (replica groups) x (site chunks of
// kWavesPerBlock*spw sites), padded to a multiple of 8 for the XCD map
static int phase_blocks(int n_sites, int groups, int xcd_map, int spw) {
    const int chunk = kWavesPerBlock * spw;
    long long nb = (long long)groups * ((n_sites + chunk - 1) / chunk);
    if (xcd_map) nb = (nb + 7) / 8 * 8;
    return (int)nb;
}

// workgroups of a phase (and its site chunks) for `spw_sites` sites per wave
template <class A>
static void set_blocks(A& a, int groups, int spw_sites) {
    const int chunk = kWavesPerBlock * spw_sites;
    a.nchunk = (a.c_end - a.c_begin + chunk - 1) / chunk;
    a.nblocks = phase_blocks(a.c_end - a.c_begin, groups, a.xcd_map, spw_sites);
}

static XyzArgs xyz_args(ether_ctx* c, int colour) {
    XyzArgs a;
    a.sx = c->d_sx;
    a.sy = c->d_sy;
    a.sz = c->d_sz;
    a.phi = c->d_phi;
    a.ell = c->d_ell;
    a.D = c->D;
    a.sp = c->d_sp;
    a.N = c->N;
    a.R = c->R;
    a.Rp = c->Rp;
    a.G = c->Rp / kWave;
    a.c_begin = c->cbeg[colour];
    a.c_end = c->cbeg[colour + 1];
    a.xcd_map = c->xcd_map;
    a.spw = c->spw;
    a.lps = c->lps;
    a.sps = kWave / c->lps;
    set_blocks(a, a.G, a.spw * a.sps);
    a.slot_of = c->d_slot_of;
    a.beta_slot = c->d_beta_f;
    a.seed = c->seed;
    a.ctr = c->d_ctr;
    a.step_off = c->step - (c->capturing ? c->cap_step0 : c->dev_step);
    a.pass = 0;
    a.frac = 1.0f;
    a.fuse = 0;
    a.plane_bytes = (uint32_t)((size_t)c->N * c->Rp * sizeof(float));
    a.acc = c->d_acc;
    a.n_classes = c->n_classes;
    a.jclass = c->d_jclass;
    a.nsp = c->nsp;
    a.zvec = c->d_zvec;
    a.dsia = c->d_dsia;
    a.dphi = (float)c->dphi;
    a.zee = c->zee;
    a.sia = c->sia;
    a.meas = 0;
    a.meas_lower = 0;
    a.w = c->d_w;
    a.macc = c->d_macc;
    a.mrec = nullptr;
    a.mrec_stride = 0;
    a.mrec_pp = 0;
    a.mrec_phase = 0;
    for (int k = 0; k < 4; ++k) a.mscale[k] = c->mscale[k];
    return a;
}

static IsingArgs ising_args(ether_ctx* c, int colour) {
    IsingArgs a;
    a.s = c->d_s;
    a.ell = c->d_ell;
    a.D = c->D;
    a.sp = c->d_sp;
    a.N = c->N;
    a.R = c->R;
    a.Rp = c->Rp;
    a.G4 = (c->Rp + 4 * kWave - 1) / (4 * kWave);
    a.c_begin = c->cbeg[colour];
    a.c_end = c->cbeg[colour + 1];
    a.xcd_map = c->xcd_map;
    a.spw = c->spw;
    a.lps_shift = lanes_shift_of(c->lps);   // a power of two for Ising
    set_blocks(a, a.G4, a.spw << (6 - a.lps_shift));
    a.table_in_lds = ((long long)c->R * c->ising_table <= kIsingLdsTable) ? 1 : 0;
    a.slot_of = c->d_slot_of;
    a.seed = c->seed;
    a.ctr = c->d_ctr;
    a.step_off = c->step - (c->capturing ? c->cap_step0 : c->dev_step);
    a.acc = c->d_acc;
    a.n_classes = c->n_classes;
    a.stride = c->d_stride;
    a.offset = c->ising_offset;
    a.table_per_slot = c->ising_table;
    a.P = c->ising_P;
    a.thr = c->d_thr;
    return a;
}

static MeasArgs meas_args(ether_ctx* c) {
    MeasArgs a;
    a.sx = c->d_sx;
    a.sy = c->d_sy;
    a.sz = c->d_sz;
    a.s = c->d_s;
    a.ell = c->d_ell;
    a.D = c->D;
    a.sp = c->d_sp;
    a.w = c->d_w;
    a.N = c->N;
    a.R = c->R;
    a.Rp = c->Rp;
    a.G = c->Rp / kWave;
    a.tile_sites = c->meas_sites;
    a.tiles = c->meas_tiles;
    a.rec = c->meas_rec;
    a.n_classes = c->n_classes;
    a.jclass = c->d_jclass;
    a.nsp = c->nsp;
    a.zvec = c->d_zvec;
    a.dsia = c->d_dsia;
    a.ell_m = c->d_ell_m ? c->d_ell_m : c->d_ell;
    a.ell_m_n = c->d_ell_m_n;
    a.D_m = c->D_m;
    a.bond_factor = c->bond_factor;
    a.plane_bytes = (uint32_t)((size_t)c->Rp * c->N * sizeof(float));
    a.spiral = c->d_spiral;
    a.spiral_cap = c->model == ETHER_MODEL_ISING ? 0 : c->spiral_cap;
    a.partial = c->d_partial;
    return a;
}

// --------------------------------------------------------------------------
// kernel timing (HIP events on the context stream)
// --------------------------------------------------------------------------
// An event pair brackets `count` back-to-back launches of one kind (the
// colour phases of a sweep share one pair, so the events' own cost is paid
// once per sweep rather than once per phase).
static int ev_begin(ether_ctx* c, int kind, int count = 1) {
    if (!c->timing) return -1;
    if (c->ev_used == c->ev_pool.size()) {
        hipEvent_t a, b;
        if (hipEventCreate(&a) != hipSuccess) return -1;
        if (hipEventCreate(&b) != hipSuccess) {
            (void)hipEventDestroy(a);
            return -1;
        }
        c->ev_pool.push_back({a, b});
        c->ev_kind.push_back(kind);
        c->ev_count.push_back(count);
    }
    const int k = (int)c->ev_used++;
    c->ev_kind[k] = kind;
    c->ev_count[k] = count;
    (void)hipEventRecord(c->ev_pool[k].first, cst(c));
    return k;
}

static void ev_end(ether_ctx* c, int k) {
    if (k >= 0) (void)hipEventRecord(c->ev_pool[k].second, cst(c));
}

extern "C" int ether_set_timing(ether_ctx* c, int32_t on) {
    if (!c) return fail(nullptr, ETHER_ERR_ARG, "null context");
    HIPCHK(c, hipStreamSynchronize(cst(c)));
    c->timing = on != 0;
    c->ev_used = 0;
    return ETHER_OK;
}

extern "C" int ether_get_timing(ether_ctx* c, int32_t kind, double* total_ms, int64_t* launches) {
    if (!c) return fail(nullptr, ETHER_ERR_ARG, "null context");
    HIPCHK(c, hipStreamSynchronize(cst(c)));
    double tot = 0.0;
    int64_t n = 0;
    for (size_t k = 0; k < c->ev_used; ++k) {
        if (c->ev_kind[k] != kind) continue;
        float ms = 0.f;
        HIPCHK(c, hipEventElapsedTime(&ms, c->ev_pool[k].first, c->ev_pool[k].second));
        tot += ms;
        n += c->ev_count[k];
    }
    if (total_ms) *total_ms = tot;
    if (launches) *launches = n;
    return ETHER_OK;
}

extern "C" int ether_synchronize(ether_ctx* c) {
    if (!c) return fail(nullptr, ETHER_ERR_ARG, "null context");
    HIPCHK(c, hipStreamSynchronize(cst(c)));
    return ETHER_OK;
}

// --------------------------------------------------------------------------
// hot path
// --------------------------------------------------------------------------
// One sweep: a Metropolis phase per colour.  fuse_frac > 0 (XYZ only) makes
// every phase first reflect each of its sites with probability fuse_frac.
// While a batch records, every context emits one launch per colour up to the
// batch's largest colour count (empty ones included), so that the recorded
// sequences of all lattices line up launch by launch.
static int phase_count(const ether_ctx* c) {
    return g_recorder ? std::max(c->ncol, c->rec_ncol) : c->ncol;
}
static bool phase_empty(const ether_ctx* c, int q) {
    return q >= c->ncol || c->cbeg[q + 1] == c->cbeg[q];
}

static bool resident_path(const ether_ctx* c);
static int enqueue_resident(ether_ctx* c, int n_sweeps, bool meas);

// meas: the phases also book the fused measurement that the next
// enqueue_measure(..., fused = true) completes.
// The measuring sweep's shape with per-workgroup records, and the workgroups
// of each of its phases (the most any colour needs at that shape).
static int meas_record_spw(const ether_ctx* c) { return std::max(c->spw, kMeasRecordSpw); }
static int meas_record_stride(const ether_ctx* c) {
    const int spw = meas_record_spw(c) * (kWave / c->lps);
    int n = 0;
    for (int q = 0; q < c->ncol; ++q)
        n = std::max(n, phase_blocks(c->cbeg[q + 1] - c->cbeg[q], c->Rp / kWave, c->xcd_map, spw));
    return n;
}

static int enqueue_sweep(ether_ctx* c, double fuse_frac = 0.0, bool meas = false) {
    int launches = 0;
    for (int q = 0; q < phase_count(c); ++q)
        if (!phase_empty(c, q) || g_recorder) ++launches;
    const int ev = ev_begin(c, meas ? ETHER_TIME_SWEEP_MEAS : ETHER_TIME_SWEEP, launches);
    for (int q = 0; q < phase_count(c); ++q) {
        const bool empty = phase_empty(c, q);
        if (empty && !g_recorder) continue;
        const int qq = std::min(q, c->ncol - 1);
        if (c->model == ETHER_MODEL_ISING) {
            IsingArgs a = ising_args(c, qq);
            if (empty) a.c_begin = a.c_end = 0;
            launch_ising_sweep(a, cst(c));
        } else {
            XyzArgs a = xyz_args(c, qq);
            if (empty) a.c_begin = a.c_end = 0;
            if (fuse_frac > 0.0) {
                a.fuse = 1;
                a.frac = (float)fuse_frac;
            }
            if (meas) {
                a.meas = 1;
                a.meas_lower = empty ? 0 : c->meas_lower[qq];
                if (c->meas_records) {
                    // one record per workgroup: the run's shape (at least
                    // kMeasRecordSpw site steps per wave) with the same
                    // number of workgroups in every phase, so that each
                    // measuring sweep rewrites every record the finalize reads
                    a.spw = meas_record_spw(c);
                    set_blocks(a, a.G, a.spw * a.sps);
                    a.nblocks = meas_record_stride(c);
                    if (a.nblocks > c->mrec_cap)
                        return fail(c, ETHER_ERR_STATE, "measurement records exceed their allocation");
                    a.mrec = c->d_mrec;
                    a.mrec_stride = a.nblocks;
                    a.mrec_pp = c->mrec_pp;
                    a.mrec_phase = qq;
                } else {
                    // a workgroup books its sums with one atomic per (replica,
                    // word), so the phase's atomics scale with its workgroups:
                    // the longest walk per wave keeps them few (the spins do not
                    // depend on the launch shape).  Several sites per wave
                    // (few replicas) keep the context's walk: there a
                    // workgroup books fewer replicas, and the long walk cost
                    // more in latency than it saved in atomics (62 simple-cubic
                    // Heisenberg lattices of 8..32 replicas batched: 0.754 ->
                    // 0.715 ms per MCS)
                    a.spw = c->lps < kWave ? c->spw : kMaxSitesPerWave;
                    set_blocks(a, a.G, a.spw * a.sps);
                }
            }
            launch_xyz_sweep(a, c->angle, c->zee, c->sia, cst(c));
        }
    }
    ev_end(c, ev);
    HIPCHK(c, hipGetLastError());
    c->step++;
    return ETHER_OK;
}

static int enqueue_ovrr(ether_ctx* c, double passes) {
    int full = (int)std::floor(passes);
    double frac = passes - full;
    int pass = 0;
    for (int p = 0; p < full + (frac > 0.0 ? 1 : 0); ++p, ++pass) {
        for (int q = 0; q < phase_count(c); ++q) {
            const bool empty = phase_empty(c, q);
            if (empty && !g_recorder) continue;
            XyzArgs a = xyz_args(c, std::min(q, c->ncol - 1));
            if (empty) a.c_begin = a.c_end = 0;
            a.lps = kWave;   // the overrelaxation kernel updates one site per wave
            a.sps = 1;
            set_blocks(a, a.G, a.spw);
            a.pass = (uint32_t)pass;
            a.frac = (p < full) ? 1.0f : (float)frac;
            const int ev = ev_begin(c, ETHER_TIME_OVRR);
            launch_xyz_ovrr(a, c->angle, c->zee, cst(c));
            ev_end(c, ev);
        }
    }
    HIPCHK(c, hipGetLastError());
    return ETHER_OK;
}

// fused: the record was booked by the sweep just enqueued (enqueue_sweep with
// meas); only the finalize runs.  While a batch records, an empty measurement
// launch keeps the launch sequence aligned with lattices that measure apart.
// booked: what precedes the finalize.  kBookedPass: nothing (a measurement
// pass runs here); kBookedXyz: the sweep's fixed-point sums (d_macc);
// kBookedResident: the resident Ising kernel's record in tile 0.
enum Booked : int { kBookedPass = 0, kBookedXyz = 1, kBookedResident = 2 };
static int enqueue_measure(ether_ctx* c, int sample, int booked = kBookedPass) {
    const int ev = ev_begin(c, ETHER_TIME_MEASURE);
    if (booked == kBookedPass) {
        launch_measure(meas_args(c), c->model, c->zee, c->sia, cst(c));
    } else if (g_recorder) {
        MeasArgs m = meas_args(c);
        m.tiles = 0;
        launch_measure(m, c->model, c->zee, c->sia, cst(c));
    }
    FinArgs f;
    f.partial = c->d_partial;
    f.tiles = booked == kBookedResident ? 1 : c->meas_tiles;
    f.rec = c->meas_rec;
    f.Rp = c->Rp;
    f.R = c->R;
    f.N = c->N;
    f.model = c->model;
    f.n_classes = c->n_classes;
    f.ising_j = c->d_ising_j;
    f.nsp = c->nsp;
    f.ising_zc = c->d_ising_zc;
    f.sample = sample;
    f.pt_etot = (c->compat & ETHER_COMPAT_PT_ETOT) ? 1 : 0;
    f.E_pt = c->d_Ept;
    f.parts = c->d_parts;
    f.M = c->d_M;
    f.bonds = c->d_bonds;
    f.moments = c->d_mom;
    f.slot_of = c->d_slot_of;
    f.spiral_count = c->spiral_cap ? c->spiral_count : 0.0;
    f.psi = c->d_psi;
    f.mom_sp = c->d_mom_sp;
    f.macc = (booked == kBookedXyz && !c->meas_records) ? c->d_macc : nullptr;
    f.mrec = (booked == kBookedXyz && c->meas_records) ? c->d_mrec : nullptr;
    f.mrec_stride = c->meas_records ? meas_record_stride(c) : 0;
    f.mrec_phases = c->ncol;
    f.mrec_groups = c->Rp / kWave;
    f.mrec_pp = c->mrec_pp;
    for (int k = 0; k < 4; ++k) f.mscale[k] = c->mscale[k];
    launch_finalize(f, cst(c));
    ev_end(c, ev);
    HIPCHK(c, hipGetLastError());
    return ETHER_OK;
}

static int enqueue_pt(ether_ctx* c, int parity, int booked = kBookedPass) {
    if (c->n_groups <= 0) return ETHER_OK;
    TRY(enqueue_measure(c, 0, booked));
    PtArgs p;
    p.n_groups = c->n_groups;
    p.gstart = c->d_gstart;
    // inside a captured graph the parity comes from the device swap counter
    p.parity = c->capturing ? -1 : (parity & 1);
    p.rep_at = c->d_rep_at;
    p.slot_of = c->d_slot_of;
    p.beta_slot = c->d_beta;
    p.E_pt = c->d_Ept;
    p.seed = c->seed;
    p.ctr = c->d_ctr;
    p.pt_off = c->pt_count - (c->capturing ? c->cap_pt0 : c->dev_pt);
    c->pt_count++;
    p.acc = c->d_pt_acc;
    p.att = c->d_pt_att;
    p.trips = c->d_trips;
    p.trip_state = c->d_trip_state;
    p.n_accepted = c->d_nacc;
    const int ev = ev_begin(c, ETHER_TIME_PT);
    launch_pt_swap(p, cst(c));
    ev_end(c, ev);
    HIPCHK(c, hipGetLastError());
    return ETHER_OK;
}

extern "C" int ether_init_spins(ether_ctx* c, int32_t mode) {
    TRY(need_committed(c));
    if (mode != ETHER_INIT_FRESH && mode != ETHER_INIT_RANDOM)
        return fail(c, ETHER_ERR_ARG, "ether_init_spins: mode must be FRESH or RANDOM "
                                      "(set a basis state with ether_set_spins)");
    InitArgs a;
    a.sx = c->d_sx;
    a.sy = c->d_sy;
    a.sz = c->d_sz;
    a.phi = c->d_phi;
    a.s = c->d_s;
    a.N = c->N;
    a.R = c->R;
    a.Rp = c->Rp;
    a.model = c->model;
    a.angle = c->angle;
    a.mode = mode;
    a.dphi = (float)c->dphi;
    a.seed = c->seed;
    a.step = c->init_count++;
    launch_init(a, cst(c));
    HIPCHK(c, hipGetLastError());
    return ETHER_OK;
}

extern "C" int ether_sweep(ether_ctx* c, int32_t n_mcs, double* accept) {
    TRY(need_betas(c));
    if (n_mcs < 0) return fail(c, ETHER_ERR_ARG, "ether_sweep: negative n_mcs");
    std::vector<unsigned long long> before;
    if (resident_path(c)) {
        // all but the last sweep in one launch, the last one on its own when
        // its acceptance is asked for
        const int head = accept ? n_mcs - 1 : n_mcs;
        if (head > 0) TRY(enqueue_resident(c, head, false));
        if (accept && n_mcs > 0) {
            TRY(copy_acc_async(c, before));
            TRY(enqueue_resident(c, 1, false));
        }
    } else {
        for (int k = 0; k < n_mcs; ++k) {
            if (accept && k == n_mcs - 1) TRY(copy_acc_async(c, before));
            TRY(enqueue_sweep(c));
        }
    }
    if (accept && n_mcs > 0) {
        std::vector<unsigned long long> after;
        TRY(copy_acc_async(c, after));
        HIPCHK(c, hipStreamSynchronize(cst(c)));
        for (int r = 0; r < c->R; ++r)
            accept[r] = (double)(acc_total(after, c->R, r) - acc_total(before, c->R, r)) / c->N;
    }
    return ETHER_OK;
}

// --------------------------------------------------------------------------
// launch shape of the sweep phases and its autotuner
// --------------------------------------------------------------------------
extern "C" int ether_set_launch(ether_ctx* c, int32_t sites_per_wave, int32_t xcd_map) {
    if (!c) return fail(nullptr, ETHER_ERR_ARG, "null context");
    if (sites_per_wave < 1 || sites_per_wave > kMaxSitesPerWave || xcd_map < 0 || xcd_map > 2)
        return fail(c, ETHER_ERR_ARG, "ether_set_launch: sites_per_wave must be 1..16, xcd_map 0, 1 or 2");
    if (c->spw != sites_per_wave || c->xcd_map != xcd_map) drop_graphs(c);   // baked into graphs
    c->spw = sites_per_wave;
    c->xcd_map = xcd_map;
    return ETHER_OK;
}

// Times `reps` sweeps for every (sites per wave, XCD map) candidate on the
// context's own lattice, replicas and couplings, keeps the fastest, and puts
// the spins, acceptance counters and sweep counter back as they were.
extern "C" int ether_autotune(ether_ctx* c, int32_t reps, int32_t* sites_per_wave, int32_t* xcd_map,
                              double* table_ms) {
    TRY(need_betas(c));
    if (reps <= 0) return fail(c, ETHER_ERR_ARG, "ether_autotune: reps must be positive");
    static const int kSpw[] = {1, 2, 4, 8, 16};
    const int n_spw = (int)(sizeof(kSpw) / sizeof(kSpw[0]));
    constexpr int kXcdMaps = 3;
    const size_t RN = (size_t)c->Rp * c->N;
    std::vector<std::pair<void*, size_t>> planes;
    if (c->model == ETHER_MODEL_ISING) {
        planes.push_back({c->d_s, RN});
    } else {
        planes.push_back({c->d_sx, RN * 4});
        planes.push_back({c->d_sy, RN * 4});
        planes.push_back({c->d_sz, RN * 4});
        if (c->angle) planes.push_back({c->d_phi, RN * 4});
    }
    planes.push_back({c->d_acc, sizeof(unsigned long long) * kAccCopies * c->R});
    size_t total = 0;
    for (auto& p : planes) total += (p.second + 255) / 256 * 256;
    char* save = nullptr;
    HIPCHK(c, hipMalloc(&save, total));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = ETHER_OK;
    const uint64_t step0 = c->step;
    const bool timing0 = c->timing;
    const int spw0 = c->spw, xcd0 = c->xcd_map;
    double best = 1e300;
    int best_spw = spw0, best_xcd = xcd0;
    c->timing = false;
    {
        size_t off = 0;
        for (auto& p : planes) {
            if (hipMemcpyAsync(save + off, p.first, p.second, hipMemcpyDeviceToDevice, cst(c)) !=
                hipSuccess)
                rc = fail(c, ETHER_ERR_HIP, "ether_autotune: save copy failed");
            off += (p.second + 255) / 256 * 256;
        }
    }
    if (!rc && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
        rc = fail(c, ETHER_ERR_HIP, "ether_autotune: event creation failed");
    // two rounds over the candidates, each keeping its faster round: a
    // candidate's time then does not hinge on one round's clock
    std::vector<double> per_ms(kXcdMaps * n_spw, 1e300);
    for (int round = 0; round < 2 && !rc; ++round)
        for (int i = 0; !rc && i < kXcdMaps * n_spw; ++i) {
            c->spw = kSpw[i % n_spw];
            c->xcd_map = i / n_spw;
            if ((rc = enqueue_sweep(c))) break;   // warm-up of this shape
            (void)hipEventRecord(e0, cst(c));
            for (int k = 0; k < reps && !rc; ++k) rc = enqueue_sweep(c);
            (void)hipEventRecord(e1, cst(c));
            if (rc) break;
            if (hipEventSynchronize(e1) != hipSuccess) {
                rc = fail(c, ETHER_ERR_HIP, "ether_autotune: sweep failed");
                break;
            }
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, e0, e1);
            per_ms[i] = std::min(per_ms[i], (double)ms / reps);
        }
    // the fastest candidate; among those within 1% of it, the longest walk
    // per wave (fewer workgroups, so fewer atomics at the end of a phase),
    // then the faster: shapes a hair apart do not trade places from run to run
    for (int i = 0; i < kXcdMaps * n_spw; ++i) best = std::min(best, per_ms[i]);
    int pick = -1;
    for (int i = 0; i < kXcdMaps * n_spw; ++i) {
        if (table_ms) table_ms[i] = per_ms[i];
        if (per_ms[i] > 1.01 * best) continue;
        if (pick < 0 || kSpw[i % n_spw] > kSpw[pick % n_spw] ||
            (kSpw[i % n_spw] == kSpw[pick % n_spw] && per_ms[i] < per_ms[pick]))
            pick = i;
    }
    if (!rc && pick >= 0) {
        best_spw = kSpw[pick % n_spw];
        best_xcd = pick / n_spw;
    }
    {
        size_t off = 0;
        for (auto& p : planes) {
            (void)hipMemcpyAsync(p.first, save + off, p.second, hipMemcpyDeviceToDevice, cst(c));
            off += (p.second + 255) / 256 * 256;
        }
    }
    (void)hipStreamSynchronize(cst(c));
    (void)hipFree(save);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    c->step = step0;
    c->timing = timing0;
    c->spw = spw0;
    c->xcd_map = xcd0;
    if (rc) return rc;
    TRY(ether_set_launch(c, best_spw, best_xcd));
    if (sites_per_wave) *sites_per_wave = best_spw;
    if (xcd_map) *xcd_map = best_xcd;
    return ETHER_OK;
}

extern "C" int ether_overrelax(ether_ctx* c, double passes) {
    TRY(need_committed(c));
    if (c->model != ETHER_MODEL_XYZ) return fail(c, ETHER_ERR_STATE, "overrelaxation needs XYZ spins");
    if (c->sia)
        return fail(c, ETHER_ERR_STATE,
                    "overrelaxation does not conserve the single-ion anisotropy energy");
    if (!(passes >= 0.0)) return fail(c, ETHER_ERR_ARG, "ether_overrelax: negative passes");
    return enqueue_ovrr(c, passes);
}

extern "C" int ether_measure(ether_ctx* c, int32_t sample, double* E, double* M) {
    TRY(need_committed(c));
    TRY(enqueue_measure(c, sample ? 1 : 0));
    if (E || M) {
        std::vector<double> parts(4 * (size_t)c->R);
        HIPCHK(c, hipMemcpyAsync(parts.data(), c->d_parts, sizeof(double) * 4 * c->R,
                                 hipMemcpyDeviceToHost, cst(c)));
        if (M)
            HIPCHK(c, hipMemcpyAsync(M, c->d_M, sizeof(double) * 3 * c->R, hipMemcpyDeviceToHost,
                                     cst(c)));
        HIPCHK(c, hipStreamSynchronize(cst(c)));
        if (E)
            for (int r = 0; r < c->R; ++r) E[r] = parts[4 * r] + parts[4 * r + 1] + parts[4 * r + 2];
    }
    return ETHER_OK;
}

extern "C" int ether_get_energy_parts(ether_ctx* c, double* parts, int64_t* bonds, int32_t* n_classes) {
    TRY(need_committed(c));
    if (parts)
        HIPCHK(c, hipMemcpyAsync(parts, c->d_parts, sizeof(double) * 4 * c->R, hipMemcpyDeviceToHost,
                                 cst(c)));
    if (n_classes) *n_classes = c->n_classes;
    if (bonds && c->model == ETHER_MODEL_ISING) {
        std::vector<long long> b((size_t)c->R * (c->n_classes + 1));
        HIPCHK(c, hipMemcpyAsync(b.data(), c->d_bonds, sizeof(long long) * b.size(),
                                 hipMemcpyDeviceToHost, cst(c)));
        HIPCHK(c, hipStreamSynchronize(cst(c)));
        for (int r = 0; r < c->R; ++r)
            for (int k = 1; k <= c->n_classes; ++k)
                bonds[(size_t)r * c->n_classes + (k - 1)] = b[(size_t)r * (c->n_classes + 1) + k];
    }
    HIPCHK(c, hipStreamSynchronize(cst(c)));
    return ETHER_OK;
}

extern "C" int ether_read_moments(ether_ctx* c, double* moments) {
    TRY(need_committed(c));
    if (!moments) return fail(c, ETHER_ERR_ARG, "ether_read_moments: null pointer");
    std::vector<unsigned long long> acc;
    HIPCHK(c, hipMemcpyAsync(moments, c->d_mom, sizeof(double) * ETHER_MOM_COUNT * c->R,
                             hipMemcpyDeviceToHost, cst(c)));
    TRY(copy_acc_async(c, acc));
    HIPCHK(c, hipStreamSynchronize(cst(c)));
    for (int r = 0; r < c->R; ++r)
        moments[ETHER_MOM_COUNT * r + ETHER_MOM_ACC] = (double)acc_total(acc, c->R, r) / c->N;
    return ETHER_OK;
}

extern "C" int ether_reset_moments(ether_ctx* c) {
    TRY(need_committed(c));
    HIPCHK(c, hipMemsetAsync(c->d_mom, 0, sizeof(double) * ETHER_MOM_COUNT * c->R, cst(c)));
    if (c->d_mom_sp)
        HIPCHK(c, hipMemsetAsync(c->d_mom_sp, 0, sizeof(double) * 3 * c->R, cst(c)));
    HIPCHK(c, hipMemsetAsync(c->d_acc, 0, sizeof(unsigned long long) * kAccCopies * c->R, cst(c)));
    return ETHER_OK;
}

extern "C" int ether_set_spiral(ether_ctx* c, int32_t capacity, const int32_t* partners) {
    if (!c) return fail(c, ETHER_ERR_ARG, "ether_set_spiral: null context");
    if (c->committed) return fail(c, ETHER_ERR_STATE, "ether_set_spiral must precede ether_commit");
    if (capacity == 0) {
        c->spiral_cap = 0;
        c->spiral_count = 0.0;
        c->spiral_rows.clear();
        return ETHER_OK;
    }
    if (c->model != ETHER_MODEL_XYZ)
        return fail(c, ETHER_ERR_ARG, "the spiral-state parameter needs vector spins (ssp with ISING)");
    if (capacity < 0 || capacity > 4096 || !partners)
        return fail(c, ETHER_ERR_ARG, "ether_set_spiral: capacity out of range or null partners");
    const int N = c->N;
    long long count = 0;
    for (int i = 0; i < N; ++i) {
        bool ended = false;
        for (int k = 0; k < capacity; ++k) {
            const int j = partners[(size_t)i * capacity + k];
            if (j < 0) {
                ended = true;
                continue;
            }
            if (ended || j >= N || j == i)
                return fail(c, ETHER_ERR_ARG, "ether_set_spiral: partner row " + std::to_string(i) +
                                                  " has an entry out of range, a self pair or a gap");
            ++count;
        }
    }
    if (count == 0)
        return fail(c, ETHER_ERR_ARG, "ether_set_spiral: no partner pairs (spiral_count = 0)");
    c->spiral_cap = capacity;
    c->spiral_count = (double)count;
    c->spiral_rows.assign(partners, partners + (size_t)N * capacity);
    return ETHER_OK;
}

extern "C" int ether_read_spiral(ether_ctx* c, double* moments, double* psi) {
    TRY(need_committed(c));
    if (!c->spiral_cap) return fail(c, ETHER_ERR_STATE, "ether_read_spiral: no spiral partners set");
    if (moments)
        HIPCHK(c, hipMemcpyAsync(moments, c->d_mom_sp, sizeof(double) * 3 * c->R, hipMemcpyDeviceToHost,
                                 cst(c)));
    if (psi)
        HIPCHK(c, hipMemcpyAsync(psi, c->d_psi, sizeof(double) * c->R, hipMemcpyDeviceToHost, cst(c)));
    HIPCHK(c, hipStreamSynchronize(cst(c)));
    return ETHER_OK;
}

extern "C" int ether_pt_swap(ether_ctx* c, int32_t parity, int32_t* n_accepted) {
    TRY(need_betas(c));
    if (c->n_groups <= 0) return fail(c, ETHER_ERR_STATE, "ether_pt_swap: no ladders set");
    HIPCHK(c, hipMemsetAsync(c->d_nacc, 0, sizeof(int), cst(c)));
    TRY(enqueue_pt(c, parity));
    if (n_accepted) {
        HIPCHK(c, hipMemcpyAsync(n_accepted, c->d_nacc, sizeof(int), hipMemcpyDeviceToHost, cst(c)));
        HIPCHK(c, hipStreamSynchronize(cst(c)));
    }
    return ETHER_OK;
}

extern "C" int ether_pt_state(ether_ctx* c, int32_t* slot_of, int64_t* acc, int64_t* att,
                              int64_t* trips) {
    TRY(need_committed(c));
    if (slot_of)
        HIPCHK(c, hipMemcpyAsync(slot_of, c->d_slot_of, sizeof(int) * c->R, hipMemcpyDeviceToHost,
                                 cst(c)));
    if (acc)
        HIPCHK(c, hipMemcpyAsync(acc, c->d_pt_acc, sizeof(int64_t) * c->R, hipMemcpyDeviceToHost,
                                 cst(c)));
    if (att)
        HIPCHK(c, hipMemcpyAsync(att, c->d_pt_att, sizeof(int64_t) * c->R, hipMemcpyDeviceToHost,
                                 cst(c)));
    if (trips)
        HIPCHK(c, hipMemcpyAsync(trips, c->d_trips, sizeof(int64_t) * c->R, hipMemcpyDeviceToHost,
                                 cst(c)));
    HIPCHK(c, hipStreamSynchronize(cst(c)));
    return ETHER_OK;
}

// One MCS of the schedule of src/ether.f90:159-203.
static int enqueue_step(ether_ctx* c, int32_t stepi, int32_t mcs_eq, int32_t coa, int ovrr_iv,
                        double ovrr_passes, int pt_iv, int* sc) {
    // src/ether.f90:163-169, then the sweep of :172.  A pass of at most one
    // reflection per site is fused into the sweep's colour phases unless the
    // split order is asked for.
    const bool ovrr = ovrr_iv && stepi % ovrr_iv == 0 && stepi % coa != 0;
    const bool fuse = ovrr && ovrr_passes <= 1.0 && !(c->compat & ETHER_COMPAT_SPLIT_OVRR);
    const bool pt = pt_iv && stepi % pt_iv == 0 && stepi % coa != 0;
    const bool sample = stepi > mcs_eq && stepi % coa == 0;
    // a measurement follows this sweep (for the swap or the sample): the
    // sweep's phases book it as they go (XYZ), instead of a pass of its own
    const bool fm = c->meas_fused_ok && !(c->compat & ETHER_COMPAT_MEASURE_PASS) &&
                    ((pt && c->n_groups > 0) || sample);
    if (ovrr && !fuse) TRY(enqueue_ovrr(c, ovrr_passes));
    TRY(enqueue_sweep(c, fuse ? ovrr_passes : 0.0, fm));
    // src/ether.f90:189-198
    if (pt) {
        ++*sc;
        TRY(enqueue_pt(c, *sc % 2, fm ? kBookedXyz : kBookedPass));
    }
    // src/ether.f90:201 -> src/observable.f90:30-31
    if (sample) TRY(enqueue_measure(c, 1, fm ? kBookedXyz : kBookedPass));
    return ETHER_OK;
}

// The resident Ising path: n_sweeps sweeps in one launch, the last one
// booking the measurement of the state it leaves when meas is set.
// While a batch records, the batch decides (g_record_resident): every
// lattice of it takes the resident path or none does, so that their launch
// sequences line up.
static thread_local bool g_record_resident = false;
static bool resident_path(const ether_ctx* c) {
    return g_recorder ? g_record_resident : (c->resident && c->res_ok);
}

static int enqueue_resident(ether_ctx* c, int n_sweeps, bool meas) {
    IsingResArgs a;
    a.s = c->d_s;
    a.ell = c->d_ell;
    a.D = c->D;
    a.sp = c->d_sp;
    a.w = c->d_w;
    a.N = c->N;
    a.R = c->R;
    a.Rp = c->Rp;
    a.ncol = c->ncol;
    for (int q = 0; q <= kMaxResColours; ++q) a.cbeg[q] = c->cbeg[std::min(q, c->ncol)];
    a.slot_of = c->d_slot_of;
    a.seed = c->seed;
    a.ctr = c->d_ctr;
    a.step_off = c->step - (c->capturing ? c->cap_step0 : c->dev_step);
    a.n_sweeps = n_sweeps;
    a.acc = c->d_acc;
    a.n_classes = c->n_classes;
    for (int k = 0; k <= kMaxIsingClasses; ++k) a.stride[k] = k <= c->n_classes ? c->ising_stride[k] : 0;
    a.offset = c->ising_offset;
    a.table_per_slot = c->ising_table;
    a.P = c->ising_P;
    // a batch runs a workgroup per replica with the table in LDS
    const bool per_rep = g_recorder || c->resident == 3 || (c->resident == 1 && c->res_rep);
    a.table_in_lds = g_recorder ? 1 : (per_rep ? c->res1_thr_lds : c->res_thr_lds);
    a.thr = c->d_thr;
    a.meas = meas ? 1 : 0;
    a.nsp = c->nsp;
    a.rec = c->meas_rec;
    a.partial = c->d_partial;
    const int ev = ev_begin(c, ETHER_TIME_RESIDENT, n_sweeps);
    if (g_recorder)
        launch_ising_resident(a, c->res1b_block, c->res1b_lds, true, cst(c));
    else
        launch_ising_resident(a, c->res_block, per_rep ? c->res1_lds : c->res_lds, per_rep, cst(c));
    ev_end(c, ev);
    HIPCHK(c, hipGetLastError());
    c->step += (uint32_t)n_sweeps;
    return ETHER_OK;
}

// Steps first..first+count-1 of the schedule.  On the resident path the
// sweeps between two measurements (a swap or a sample) are one launch, which
// books the measurement the next swap or sample reads.
static int enqueue_steps(ether_ctx* c, int32_t first, int32_t count, int32_t mcs_eq, int32_t coa,
                         int ovrr_iv, double ovrr_passes, int pt_iv, int* sc) {
    if (!resident_path(c)) {
        for (int32_t k = 0; k < count; ++k)
            TRY(enqueue_step(c, first + k, mcs_eq, coa, ovrr_iv, ovrr_passes, pt_iv, sc));
        return ETHER_OK;
    }
    int pending = 0;
    for (int32_t stepi = first; stepi < first + count; ++stepi) {
        const bool pt = pt_iv && c->n_groups > 0 && stepi % pt_iv == 0 && stepi % coa != 0;
        const bool sample = stepi > mcs_eq && stepi % coa == 0;
        ++pending;
        if (!pt && !sample) continue;
        TRY(enqueue_resident(c, pending, true));
        pending = 0;
        if (pt) {
            ++*sc;
            TRY(enqueue_pt(c, *sc % 2, kBookedResident));
        }
        if (sample) TRY(enqueue_measure(c, 1, kBookedResident));
    }
    if (pending) TRY(enqueue_resident(c, pending, false));
    return ETHER_OK;
}

static void drop_graphs(ether_ctx* c) {
    c->graph_gen++;
    (void)hipStreamSynchronize(cst(c));
    for (auto& g : c->graphs) {
        (void)hipGraphExecDestroy(g.exec);
        (void)hipGraphDestroy(g.graph);
    }
    c->graphs.clear();
}

// Capture (once) one period of P MCS, steps congruent to 1..P mod P, either
// all in equilibration (phase 0: no samples) or all in production (phase 1).
// Kernels address their random streams as device counter + offset, and the
// graph ends with a counter advance, so the same graph replays any period.
static int graph_for(ether_ctx* c, int phase, int P, int coa, int ovrr_iv, double ovrr_passes,
                     int pt_iv, int* index) {
    for (size_t q = 0; q < c->graphs.size(); ++q) {
        const auto& g = c->graphs[q];
        if (g.phase == phase && g.P == P && g.coa == coa && g.ovrr_iv == ovrr_iv &&
            g.ovrr_p == ovrr_passes && g.pt_iv == pt_iv) {
            *index = (int)q;
            return ETHER_OK;
        }
    }
    const uint32_t step0 = c->step, pt0 = c->pt_count;
    c->capturing = true;
    c->cap_step0 = step0;
    c->cap_pt0 = pt0;
    HIPCHK(c, hipStreamBeginCapture(cst(c), hipStreamCaptureModeThreadLocal));
    int rc = ETHER_OK, n_pt = 0, sc_local = 0;
    {
        // position k in the period stands for every stepi = k (mod P)
        const int32_t eq = phase ? 0 : P + 1;
        rc = enqueue_steps(c, 1, P, eq, coa, ovrr_iv, ovrr_passes, pt_iv, &sc_local);
        n_pt = sc_local;
    }
    if (rc == ETHER_OK) launch_advance_ctr(c->d_ctr, (uint32_t)P, (uint32_t)n_pt, cst(c));
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamEndCapture(cst(c), &graph);
    c->capturing = false;
    c->step = step0;
    c->pt_count = pt0;
    if (rc != ETHER_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
    }
    if (e != hipSuccess) return fail(c, ETHER_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e));
    hipGraphExec_t exec = nullptr;
    e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(graph);
        return fail(c, ETHER_ERR_HIP, std::string("graph instantiate: ") + hipGetErrorString(e));
    }
    c->graphs.push_back({phase, P, coa, ovrr_iv, pt_iv, n_pt, ovrr_passes, graph, exec});
    *index = (int)c->graphs.size() - 1;
    return ETHER_OK;
}

static int replay_graph(ether_ctx* c, int gi, int* sc) {
    const auto& g = c->graphs[gi];
    if (c->dev_step != c->step || c->dev_pt != c->pt_count || c->dev_sc != (uint32_t)*sc) {
        launch_set_ctr(c->d_ctr, c->step, c->pt_count, (uint32_t)*sc, cst(c));
        c->dev_step = c->step;
        c->dev_pt = c->pt_count;
        c->dev_sc = (uint32_t)*sc;
    }
    HIPCHK(c, hipGraphLaunch(g.exec, cst(c)));
    c->step += (uint32_t)g.P;
    c->dev_step += (uint32_t)g.P;
    c->pt_count += (uint32_t)g.n_pt;
    c->dev_pt += (uint32_t)g.n_pt;
    c->dev_sc += (uint32_t)g.n_pt;
    *sc += g.n_pt;
    return ETHER_OK;
}

extern "C" int ether_run(ether_ctx* c, int32_t first_step, int32_t n_steps, int32_t mcs_eq,
                         int32_t coa, int32_t ovrr_interval, double ovrr_passes,
                         int32_t pt_interval, int32_t* swap_count) {
    TRY(need_betas(c));
    if (coa <= 0) return fail(c, ETHER_ERR_ARG, "ether_run: coa must be positive");
    if (n_steps < 0) return fail(c, ETHER_ERR_ARG, "ether_run: negative n_steps");
    const bool do_ovrr = ovrr_interval > 0 && c->model == ETHER_MODEL_XYZ && !c->sia &&
                         ovrr_passes > 0.0;
    const bool do_pt = pt_interval > 0 && c->n_groups > 0;
    const int ovrr_iv = do_ovrr ? ovrr_interval : 0;
    const int pt_iv = do_pt ? pt_interval : 0;
    int sc = swap_count ? *swap_count : 0;
    // the schedule repeats with period P = lcm(coa, ovrr, pt); whole periods that
    // lie entirely before or after mcs_eq replay a captured graph
    auto gcd = [](long long x, long long y) { while (y) { long long t = x % y; x = y; y = t; } return x; };
    long long P = coa;
    if (ovrr_iv) P = P / gcd(P, ovrr_iv) * ovrr_iv;
    if (pt_iv) P = P / gcd(P, pt_iv) * pt_iv;
    const bool graphs_ok = c->use_graphs && !c->timing && P <= kMaxGraphPeriod;
    const int32_t last = first_step + n_steps - 1;
    int32_t stepi = first_step;
    while (stepi <= last) {
        if (graphs_ok && (stepi - 1) % P == 0 && stepi + P - 1 <= last) {
            const int phase = stepi > mcs_eq ? 1 : (stepi + P - 1 <= mcs_eq ? 0 : -1);
            if (phase >= 0) {
                int gi = -1;
                TRY(graph_for(c, phase, (int)P, coa, ovrr_iv, ovrr_passes, pt_iv, &gi));
                TRY(replay_graph(c, gi, &sc));
                stepi += (int32_t)P;
                continue;
            }
        }
        // launched one by one up to the next period boundary
        int32_t stop = stepi + 1;
        while (stop <= last && (stop - 1) % P != 0) ++stop;
        TRY(enqueue_steps(c, stepi, stop - stepi, mcs_eq, coa, ovrr_iv, ovrr_passes, pt_iv, &sc));
        stepi = stop;
    }
    if (swap_count) *swap_count = sc;
    return ETHER_OK;
}

// Captures, instantiates and uploads every graph that ether_run over the same
// steps would replay, without running anything, so that a later ether_run
// pays no capture or instantiation.
extern "C" int ether_run_prepare(ether_ctx* c, int32_t first_step, int32_t n_steps, int32_t mcs_eq,
                                 int32_t coa, int32_t ovrr_interval, double ovrr_passes,
                                 int32_t pt_interval) {
    TRY(need_betas(c));
    if (coa <= 0) return fail(c, ETHER_ERR_ARG, "ether_run_prepare: coa must be positive");
    if (n_steps < 0) return fail(c, ETHER_ERR_ARG, "ether_run_prepare: negative n_steps");
    const bool do_ovrr = ovrr_interval > 0 && c->model == ETHER_MODEL_XYZ && !c->sia &&
                         ovrr_passes > 0.0;
    const bool do_pt = pt_interval > 0 && c->n_groups > 0;
    const int ovrr_iv = do_ovrr ? ovrr_interval : 0;
    const int pt_iv = do_pt ? pt_interval : 0;
    auto gcd = [](long long x, long long y) { while (y) { long long t = x % y; x = y; y = t; } return x; };
    long long P = coa;
    if (ovrr_iv) P = P / gcd(P, ovrr_iv) * ovrr_iv;
    if (pt_iv) P = P / gcd(P, pt_iv) * pt_iv;
    if (!(c->use_graphs && !c->timing && P <= kMaxGraphPeriod)) return ETHER_OK;
    const int32_t last = first_step + n_steps - 1;
    bool want[2] = {false, false};
    for (int32_t stepi = first_step; stepi <= last;) {
        if ((stepi - 1) % P == 0 && stepi + P - 1 <= last) {
            const int phase = stepi > mcs_eq ? 1 : (stepi + P - 1 <= mcs_eq ? 0 : -1);
            if (phase >= 0) {
                want[phase] = true;
                stepi += (int32_t)P;
                continue;
            }
        }
        ++stepi;
    }
    for (int phase = 0; phase < 2; ++phase) {
        if (!want[phase]) continue;
        int gi = -1;
        TRY(graph_for(c, phase, (int)P, coa, ovrr_iv, ovrr_passes, pt_iv, &gi));
        HIPCHK(c, hipGraphUpload(c->graphs[gi].exec, cst(c)));
    }
    HIPCHK(c, hipStreamSynchronize(cst(c)));
    return ETHER_OK;
}

extern "C" int ether_set_resident(ether_ctx* c, int32_t on) {
    if (!c) return fail(nullptr, ETHER_ERR_ARG, "null context");
    if (on < 0 || on > 3) return fail(c, ETHER_ERR_ARG, "ether_set_resident: mode must be 0, 1, 2 or 3");
    drop_graphs(c);   // captured graphs hold the launches of the path they were captured on
    c->resident = on;
    return ETHER_OK;
}

extern "C" int ether_get_resident(ether_ctx* c, int32_t* kind) {
    if (!c || !kind) return fail(c, ETHER_ERR_ARG, "ether_get_resident: null argument");
    if (!c->committed) return fail(c, ETHER_ERR_STATE, "context not committed");
    if (!c->resident || !c->res_ok) *kind = 0;
    else if (c->resident == 2 || c->resident == 3) *kind = c->resident;
    else *kind = c->res_rep ? 3 : 2;
    return ETHER_OK;
}

extern "C" int ether_set_dft_resident(ether_ctx* c, int32_t on) {
    if (!c) return fail(nullptr, ETHER_ERR_ARG, "null context");
    c->dft_resident = on ? 1 : 0;
    return ETHER_OK;
}

extern "C" int ether_set_graphs(ether_ctx* c, int32_t on) {
    if (!c) return fail(nullptr, ETHER_ERR_ARG, "null context");
    c->use_graphs = on ? 1 : 0;
    return ETHER_OK;
}

// --------------------------------------------------------------------------
// batches: many lattices (one committed context each) advanced by the same
// schedule with one launch per kernel kind and template instance
// --------------------------------------------------------------------------
struct ether_batch {
    std::vector<ether_ctx*> ctx;
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    int max_ncol = 0;
    uint32_t** d_ctrs = nullptr;   // the contexts' device counters
    uint32_t* d_vals = nullptr;    // staging for counter resets
    char* d_arena = nullptr;       // argument blocks and item tables of launches issued one by one
    size_t arena_cap = 0;
    // the host image of the last upload to d_arena: it must outlive the copy,
    // which the runtime may perform after hipMemcpyAsync has returned
    std::vector<char> stage;
    hipEvent_t stage_done = nullptr;
    bool stage_pending = false;
    int use_graphs = 1;
    // every lattice sweeps with its spins resident in LDS (Ising, each
    // context's resident mode 1 or 3, and each fitting kBatchResidentLds)
    bool resident = false;
    uint64_t gen = 0;              // sum of the contexts' graph generations when graphs were captured
    struct Graph {
        int phase, P, coa, ovrr_iv, pt_iv, n_pt;
        double ovrr_p;
        hipGraph_t graph;
        hipGraphExec_t exec;
        char* d_arena;
    };
    std::vector<Graph> graphs;
};

namespace {

struct BatchedLaunch {
    int kind, variant;
    size_t args_off, items_off;
    int n_items;
    int block, lds;
};

size_t align256(size_t x) { return (x + 255) / 256 * 256; }

// Lays out the recorded launches of every context as batched launches: level
// v (the v-th recorded launch of each context) becomes one launch per (kind,
// variant) present at that level, issued in level order, so each lattice's own
// launch order is kept.  Workgroup j of a batched launch runs
// list[(j % 8) * per + j / 8] of the lattice-major list of (argument block,
// local workgroup) pairs: each XCD gets a contiguous run of one lattice.
int plan_batch(const std::vector<LaunchRecorder>& recs, std::vector<char>& host,
               std::vector<BatchedLaunch>& out, std::string& err) {
    const size_t nlev = recs[0].recs.size();
    for (const auto& r : recs)
        if (r.recs.size() != nlev) {
            err = "ether_batch_run: the lattices of a batch do not run the same schedule";
            return ETHER_ERR_STATE;
        }
    for (size_t v = 0; v < nlev; ++v) {
        const int kind = recs[0].recs[v].kind;
        std::vector<int> variants;
        for (const auto& r : recs) {
            const LaunchRecord& q = r.recs[v];
            if (q.kind != kind) {
                err = "ether_batch_run: the lattices of a batch do not run the same schedule";
                return ETHER_ERR_STATE;
            }
            if (q.units > 0 && std::find(variants.begin(), variants.end(), q.variant) == variants.end())
                variants.push_back(q.variant);
        }
        for (int var : variants) {
            const size_t aoff = align256(host.size());
            host.resize(aoff);
            std::vector<int2> list;
            std::vector<int> cost;   // LK_ISING_RESIDENT: sites of each argument block's lattice
            int na = 0, block = 0, lds = 0;
            for (const auto& r : recs) {
                const LaunchRecord& q = r.recs[v];
                if (q.units <= 0 || q.variant != var) continue;
                const size_t o = host.size();
                host.resize(o + (size_t)q.bytes);
                std::memcpy(host.data() + o, &q.u, (size_t)q.bytes);
                for (int blk = 0; blk < q.units; ++blk) list.push_back(make_int2(na, blk));
                if (kind == LK_ISING_RESIDENT) cost.push_back(q.u.r.N);
                block = std::max(block, q.block);
                lds = std::max(lds, q.lds);
                ++na;
            }
            const int n = (int)list.size();
            if (kind == LK_ISING_RESIDENT) {
                // a workgroup runs a whole replica's sweeps: the largest
                // lattices first (workgroups are dispatched in order, round
                // robin over the XCDs), the small ones fill the tail
                std::stable_sort(list.begin(), list.end(),
                                 [&](const int2& x, const int2& y) { return cost[x.x] > cost[y.x]; });
                const size_t ioff = align256(host.size());
                host.resize(ioff + sizeof(int2) * (size_t)n);
                std::memcpy(host.data() + ioff, list.data(), sizeof(int2) * (size_t)n);
                out.push_back({kind, var, aoff, ioff, n, block, lds});
                continue;
            }
            const int nb = (n + 7) / 8 * 8, per = nb / 8;
            const size_t ioff = align256(host.size());
            host.resize(ioff + sizeof(int2) * (size_t)nb);
            int2* it = reinterpret_cast<int2*>(host.data() + ioff);
            for (int j = 0; j < nb; ++j) {
                const int src = (j % 8) * per + j / 8;
                it[j] = src < n ? list[src] : make_int2(-1, 0);
            }
            out.push_back({kind, var, aoff, ioff, nb, block, lds});
        }
    }
    return ETHER_OK;
}

void issue_batch(const std::vector<BatchedLaunch>& ls, const char* d_base, hipStream_t st) {
    for (const auto& L : ls)
        launch_batch(L.kind, L.variant, d_base + L.args_off,
                     reinterpret_cast<const int2*>(d_base + L.items_off), L.n_items, L.block, L.lds, st);
}

int batch_fail(const std::string& msg, int code) {
    set_global_error(msg);
    return code;
}

}  // namespace

static void batch_drop_graphs(ether_batch* b) {
    (void)hipStreamSynchronize(b->stream);
    for (auto& g : b->graphs) {
        (void)hipGraphExecDestroy(g.exec);
        (void)hipGraphDestroy(g.graph);
        (void)hipFree(g.d_arena);
    }
    b->graphs.clear();
}

extern "C" int ether_batch_create(ether_ctx* const* ctxs, int32_t n, ether_batch** out) {
    if (!out) return batch_fail("ether_batch_create: null output", ETHER_ERR_ARG);
    *out = nullptr;
    if (!ctxs || n <= 0) return batch_fail("ether_batch_create: no contexts", ETHER_ERR_ARG);
    for (int l = 0; l < n; ++l) {
        ether_ctx* c = ctxs[l];
        if (!c || !c->committed) return batch_fail("ether_batch_create: context not committed", ETHER_ERR_STATE);
        if (c->model != ctxs[0]->model || c->sia != ctxs[0]->sia)
            return batch_fail("ether_batch_create: contexts must share the spin model and the SIA switch",
                              ETHER_ERR_ARG);
        for (int k = 0; k < l; ++k)
            if (ctxs[k] == c) return batch_fail("ether_batch_create: a context is listed twice", ETHER_ERR_ARG);
    }
    ether_batch* b = new ether_batch();
    b->ctx.assign(ctxs, ctxs + n);
    for (ether_ctx* c : b->ctx) b->max_ncol = std::max(b->max_ncol, c->ncol);
    std::vector<uint32_t*> ptrs(n);
    for (int l = 0; l < n; ++l) ptrs[l] = b->ctx[l]->d_ctr;
    hipError_t e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->done, hipEventDisableTiming);
    if (e == hipSuccess) e = hipMalloc(&b->d_ctrs, sizeof(uint32_t*) * n);
    if (e == hipSuccess) e = hipMalloc(&b->d_vals, sizeof(uint32_t) * 3 * n);
    if (e == hipSuccess)
        e = hipMemcpy(b->d_ctrs, ptrs.data(), sizeof(uint32_t*) * n, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        ether_batch_destroy(b);
        return batch_fail(std::string("ether_batch_create: ") + hipGetErrorString(e), ETHER_ERR_HIP);
    }
    *out = b;
    return ETHER_OK;
}

extern "C" void ether_batch_destroy(ether_batch* b) {
    if (!b) return;
    for (ether_ctx* c : b->ctx)
        if (c->after && c->after == b->done) (void)cst(c);   // the wait, before the event goes
    if (b->stream) batch_drop_graphs(b);
    (void)hipFree(b->d_arena);
    (void)hipFree(b->d_ctrs);
    (void)hipFree(b->d_vals);
    if (b->stage_done) {
        if (b->stage_pending) (void)hipEventSynchronize(b->stage_done);
        (void)hipEventDestroy(b->stage_done);
    }
    if (b->done) (void)hipEventDestroy(b->done);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

extern "C" int ether_batch_set_graphs(ether_batch* b, int32_t on) {
    if (!b) return batch_fail("null batch", ETHER_ERR_ARG);
    b->use_graphs = on ? 1 : 0;
    return ETHER_OK;
}

extern "C" int ether_batch_get_resident(ether_batch* b, int32_t* on) {
    if (!b || !on) return batch_fail("ether_batch_get_resident: null argument", ETHER_ERR_ARG);
    *on = b->resident ? 1 : 0;
    return ETHER_OK;
}

// Records the launches of `steps` (stepi, eq) on every context, in the mode
// the caller set up (eager, or capturing with period-relative counters).
static int batch_record(ether_batch* b, int32_t first, int32_t count, int32_t eq, int32_t coa,
                        int ovrr_iv, double ovrr_p, int pt_iv, int sc0,
                        std::vector<LaunchRecorder>& recs, int* sc_out) {
    const int n = (int)b->ctx.size();
    recs.assign(n, LaunchRecorder());
    for (int l = 0; l < n; ++l) {
        ether_ctx* c = b->ctx[l];
        int sc = sc0;
        g_recorder = &recs[l];
        g_record_resident = b->resident;
        const int rc = enqueue_steps(c, first, count, eq, coa, ovrr_iv, ovrr_p, pt_iv, &sc);
        g_recorder = nullptr;
        g_record_resident = false;
        if (rc) return batch_fail(c->err, rc);
        if (l == 0) *sc_out = sc;
        else if (sc != *sc_out)
            return batch_fail("ether_batch_run: the lattices of a batch do not run the same schedule",
                              ETHER_ERR_STATE);
    }
    return ETHER_OK;
}

static int batch_eager_step(ether_batch* b, int32_t stepi, int32_t mcs_eq, int32_t coa, int ovrr_iv,
                            double ovrr_p, int pt_iv, int* sc) {
    std::vector<LaunchRecorder> recs;
    int sc_out = *sc;
    TRY(batch_record(b, stepi, 1, mcs_eq, coa, ovrr_iv, ovrr_p, pt_iv, *sc, recs, &sc_out));
    std::vector<char> host;
    std::vector<BatchedLaunch> ls;
    std::string err;
    int rc = plan_batch(recs, host, ls, err);
    if (rc) return batch_fail(err, rc);
    if (host.size() > b->arena_cap) {
        (void)hipStreamSynchronize(b->stream);
        (void)hipFree(b->d_arena);
        b->d_arena = nullptr;
        b->arena_cap = 0;
        const size_t cap = align256(host.size() * 2);
        if (hipMalloc(&b->d_arena, cap) != hipSuccess)
            return batch_fail("ether_batch_run: argument arena allocation failed", ETHER_ERR_HIP);
        b->arena_cap = cap;
    }
    // stream-ordered after the launches that read the previous step's
    // arguments; the image lives in the batch until its copy has completed
    if (!b->stage_done && hipEventCreateWithFlags(&b->stage_done, hipEventDisableTiming) != hipSuccess)
        return batch_fail("ether_batch_run: event creation failed", ETHER_ERR_HIP);
    if (b->stage_pending && hipEventSynchronize(b->stage_done) != hipSuccess)
        return batch_fail("ether_batch_run: waiting for the previous argument upload failed", ETHER_ERR_HIP);
    b->stage_pending = false;
    b->stage.swap(host);
    if (hipMemcpyAsync(b->d_arena, b->stage.data(), b->stage.size(), hipMemcpyHostToDevice, b->stream) != hipSuccess)
        return batch_fail("ether_batch_run: argument upload failed", ETHER_ERR_HIP);
    if (hipEventRecord(b->stage_done, b->stream) != hipSuccess)
        return batch_fail("ether_batch_run: event record failed", ETHER_ERR_HIP);
    b->stage_pending = true;
    issue_batch(ls, b->d_arena, b->stream);
    if (hipGetLastError() != hipSuccess) return batch_fail("ether_batch_run: launch failed", ETHER_ERR_HIP);
    *sc = sc_out;
    return ETHER_OK;
}

// Capture (once) one period of P MCS of every lattice as one graph, as
// graph_for does for a single context; the graph ends with an advance of
// every context's device counters.
static int batch_graph_for(ether_batch* b, int phase, int P, int coa, int ovrr_iv, double ovrr_p,
                           int pt_iv, int* index) {
    for (size_t q = 0; q < b->graphs.size(); ++q) {
        const auto& g = b->graphs[q];
        if (g.phase == phase && g.P == P && g.coa == coa && g.ovrr_iv == ovrr_iv &&
            g.ovrr_p == ovrr_p && g.pt_iv == pt_iv) {
            *index = (int)q;
            return ETHER_OK;
        }
    }
    const int n = (int)b->ctx.size();
    std::vector<uint32_t> step0(n), pt0(n);
    for (int l = 0; l < n; ++l) {
        ether_ctx* c = b->ctx[l];
        step0[l] = c->step;
        pt0[l] = c->pt_count;
        c->capturing = true;
        c->cap_step0 = c->step;
        c->cap_pt0 = c->pt_count;
    }
    std::vector<LaunchRecorder> recs;
    int n_pt = 0;
    int rc = batch_record(b, 1, P, phase ? 0 : P + 1, coa, ovrr_iv, ovrr_p, pt_iv, 0, recs, &n_pt);
    for (int l = 0; l < n; ++l) {
        ether_ctx* c = b->ctx[l];
        c->capturing = false;
        c->step = step0[l];
        c->pt_count = pt0[l];
    }
    if (rc) return rc;
    std::vector<char> host;
    std::vector<BatchedLaunch> ls;
    std::string err;
    if ((rc = plan_batch(recs, host, ls, err))) return batch_fail(err, rc);
    char* arena = nullptr;
    if (hipMalloc(&arena, std::max<size_t>(host.size(), 256)) != hipSuccess)
        return batch_fail("ether_batch_run: graph argument arena allocation failed", ETHER_ERR_HIP);
    if (hipMemcpy(arena, host.data(), host.size(), hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipFree(arena);
        return batch_fail("ether_batch_run: graph argument upload failed", ETHER_ERR_HIP);
    }
    for (const auto& L : ls) grant_batch_lds(L.kind, L.variant, L.lds);
    hipError_t e = hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
        issue_batch(ls, arena, b->stream);
        launch_ctr_batch(b->d_ctrs, n, nullptr, (uint32_t)P, (uint32_t)n_pt, b->stream);
    }
    hipGraph_t graph = nullptr;
    const hipError_t e2 = hipStreamEndCapture(b->stream, &graph);
    if (e == hipSuccess) e = e2;
    hipGraphExec_t exec = nullptr;
    if (e == hipSuccess) e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        (void)hipFree(arena);
        return batch_fail(std::string("ether_batch_run: graph capture: ") + hipGetErrorString(e),
                          ETHER_ERR_HIP);
    }
    b->graphs.push_back({phase, P, coa, ovrr_iv, pt_iv, n_pt, ovrr_p, graph, exec, arena});
    *index = (int)b->graphs.size() - 1;
    return ETHER_OK;
}

static int batch_replay(ether_batch* b, int gi, int* sc) {
    const auto& g = b->graphs[gi];
    const int n = (int)b->ctx.size();
    bool reset = false;
    std::vector<uint32_t> vals(3 * (size_t)n);
    for (int l = 0; l < n; ++l) {
        ether_ctx* c = b->ctx[l];
        vals[3 * l + 0] = c->step;
        vals[3 * l + 1] = c->pt_count;
        vals[3 * l + 2] = (uint32_t)*sc;
        if (c->dev_step != c->step || c->dev_pt != c->pt_count || c->dev_sc != (uint32_t)*sc) reset = true;
    }
    if (reset) {
        // vals is freed on return: the copy must be complete before that
        if (hipMemcpyAsync(b->d_vals, vals.data(), sizeof(uint32_t) * vals.size(), hipMemcpyHostToDevice,
                           b->stream) != hipSuccess ||
            hipStreamSynchronize(b->stream) != hipSuccess)
            return batch_fail("ether_batch_run: counter upload failed", ETHER_ERR_HIP);
        launch_ctr_batch(b->d_ctrs, n, b->d_vals, 0, 0, b->stream);
        for (ether_ctx* c : b->ctx) {
            c->dev_step = c->step;
            c->dev_pt = c->pt_count;
            c->dev_sc = (uint32_t)*sc;
        }
    }
    if (hipGraphLaunch(g.exec, b->stream) != hipSuccess)
        return batch_fail("ether_batch_run: graph launch failed", ETHER_ERR_HIP);
    for (ether_ctx* c : b->ctx) {
        c->step += (uint32_t)g.P;
        c->dev_step += (uint32_t)g.P;
        c->pt_count += (uint32_t)g.n_pt;
        c->dev_pt += (uint32_t)g.n_pt;
        c->dev_sc += (uint32_t)g.n_pt;
    }
    *sc += g.n_pt;
    return ETHER_OK;
}

extern "C" int ether_batch_run(ether_batch* b, int32_t first_step, int32_t n_steps, int32_t mcs_eq,
                               int32_t coa, int32_t ovrr_interval, double ovrr_passes,
                               int32_t pt_interval, int32_t* swap_count) {
    if (!b) return batch_fail("null batch", ETHER_ERR_ARG);
    if (coa <= 0) return batch_fail("ether_batch_run: coa must be positive", ETHER_ERR_ARG);
    if (n_steps < 0) return batch_fail("ether_batch_run: negative n_steps", ETHER_ERR_ARG);
    ether_ctx* c0 = b->ctx[0];
    const bool do_ovrr = ovrr_interval > 0 && c0->model == ETHER_MODEL_XYZ && !c0->sia && ovrr_passes > 0.0;
    const bool do_pt = pt_interval > 0 && c0->n_groups > 0;
    uint64_t gen = 0;
    bool resident = c0->model == ETHER_MODEL_ISING;
    for (ether_ctx* c : b->ctx) {
        resident = resident && c->res1b_ok && (c->resident == 1 || c->resident == 3);
        int rc = need_betas(c);
        if (rc) return batch_fail(c->err, rc);
        if (c->timing) return batch_fail("ether_batch_run: kernel timing must be off in batched contexts",
                                         ETHER_ERR_STATE);
        if ((c->n_groups > 0) != (c0->n_groups > 0))
            return batch_fail("ether_batch_run: either every lattice of a batch has ladders or none",
                              ETHER_ERR_STATE);
        gen += c->graph_gen;
    }
    if (gen != b->gen || resident != b->resident) {   // a context changed what captured graphs bake in
        batch_drop_graphs(b);
        b->gen = gen;
        b->resident = resident;
    }
    for (ether_ctx* c : b->ctx) {
        (void)hipStreamSynchronize(cst(c));
        c->rec_ncol = b->max_ncol;
    }
    const int ovrr_iv = do_ovrr ? ovrr_interval : 0;
    const int pt_iv = do_pt ? pt_interval : 0;
    int sc = swap_count ? *swap_count : 0;
    auto gcd = [](long long x, long long y) { while (y) { long long t = x % y; x = y; y = t; } return x; };
    long long P = coa;
    if (ovrr_iv) P = P / gcd(P, ovrr_iv) * ovrr_iv;
    if (pt_iv) P = P / gcd(P, pt_iv) * pt_iv;
    const bool graphs_ok = b->use_graphs && P <= kMaxGraphPeriod;
    const int32_t last = first_step + n_steps - 1;
    int32_t stepi = first_step;
    int rc = ETHER_OK;
    while (stepi <= last && rc == ETHER_OK) {
        if (graphs_ok && (stepi - 1) % P == 0 && stepi + P - 1 <= last) {
            const int phase = stepi > mcs_eq ? 1 : (stepi + P - 1 <= mcs_eq ? 0 : -1);
            if (phase >= 0) {
                int gi = -1;
                rc = batch_graph_for(b, phase, (int)P, coa, ovrr_iv, ovrr_passes, pt_iv, &gi);
                if (rc == ETHER_OK) rc = batch_replay(b, gi, &sc);
                stepi += (int32_t)P;
                continue;
            }
        }
        rc = batch_eager_step(b, stepi, mcs_eq, coa, ovrr_iv, ovrr_passes, pt_iv, &sc);
        ++stepi;
    }
    // later work on each context's own stream waits for the batch
    (void)hipEventRecord(b->done, b->stream);
    for (ether_ctx* c : b->ctx) {
        c->rec_ncol = 0;
        c->after = b->done;
    }
    if (rc) return rc;
    if (swap_count) *swap_count = sc;
    return ETHER_OK;
}

// --------------------------------------------------------------------------
// spins in and out
// --------------------------------------------------------------------------
extern "C" int ether_get_spins(ether_ctx* c, int32_t replica, double* S, double* phi) {
    TRY(need_committed(c));
    if (replica < 0 || replica >= c->R || !S) return fail(c, ETHER_ERR_ARG, "ether_get_spins: bad args");
    const int N = c->N;
    float* buf = c->d_scratch_h;   // 4N floats: x, y, z, phi planes of one replica
    if (c->model == ETHER_MODEL_ISING) {
        launch_extract(nullptr, c->d_s, N, c->Rp, replica, buf, cst(c));
    } else {
        launch_extract(c->d_sx, nullptr, N, c->Rp, replica, buf, cst(c));
        launch_extract(c->d_sy, nullptr, N, c->Rp, replica, buf + N, cst(c));
        launch_extract(c->d_sz, nullptr, N, c->Rp, replica, buf + 2 * (size_t)N, cst(c));
        if (c->angle) launch_extract(c->d_phi, nullptr, N, c->Rp, replica, buf + 3 * (size_t)N, cst(c));
    }
    HIPCHK(c, hipGetLastError());
    std::vector<float> h(4 * (size_t)N, 0.f);
    HIPCHK(c, hipMemcpyAsync(h.data(), buf, sizeof(float) * 4 * N, hipMemcpyDeviceToHost, cst(c)));
    HIPCHK(c, hipStreamSynchronize(cst(c)));
    for (int i = 0; i < N; ++i) {
        const int s = c->slot_of_site[i];
        if (c->model == ETHER_MODEL_ISING) {
            S[3 * i] = 0.0;
            S[3 * i + 1] = 0.0;
            S[3 * i + 2] = (double)h[s];
            if (phi) phi[i] = 0.0;
        } else {
            S[3 * i] = h[s];
            S[3 * i + 1] = h[N + s];
            S[3 * i + 2] = h[2 * (size_t)N + s];
            if (phi) phi[i] = c->angle ? (double)h[3 * (size_t)N + s] : 0.0;
        }
    }
    return ETHER_OK;
}

extern "C" int ether_set_spins(ether_ctx* c, int32_t replica, const double* S, const double* phi) {
    TRY(need_committed(c));
    if (replica < -1 || replica >= c->R || !S)
        return fail(c, ETHER_ERR_ARG, "ether_set_spins: bad args (replica -1 sets all)");
    const int N = c->N;
    const int r0 = replica < 0 ? 0 : replica, r1 = replica < 0 ? c->R : replica + 1;
    std::vector<float> h(4 * (size_t)N, 0.f);
    for (int i = 0; i < N; ++i) {
        const int s = c->slot_of_site[i];
        h[s] = (float)S[3 * i];
        h[N + s] = (float)S[3 * i + 1];
        h[2 * (size_t)N + s] = (float)S[3 * i + 2];
        h[3 * (size_t)N + s] = phi ? (float)phi[i] : (float)std::atan2(S[3 * i + 1], S[3 * i]);
    }
    float* buf = c->d_scratch_h;
    HIPCHK(c, hipMemcpyAsync(buf, h.data(), sizeof(float) * 4 * N, hipMemcpyHostToDevice, cst(c)));
    if (c->model == ETHER_MODEL_ISING) {
        launch_insert(nullptr, c->d_s, N, c->Rp, r0, r1, buf + 2 * (size_t)N, cst(c));
    } else {
        launch_insert(c->d_sx, nullptr, N, c->Rp, r0, r1, buf, cst(c));
        launch_insert(c->d_sy, nullptr, N, c->Rp, r0, r1, buf + N, cst(c));
        launch_insert(c->d_sz, nullptr, N, c->Rp, r0, r1, buf + 2 * (size_t)N, cst(c));
        if (c->angle) launch_insert(c->d_phi, nullptr, N, c->Rp, r0, r1, buf + 3 * (size_t)N, cst(c));
    }
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(cst(c)));
    return ETHER_OK;
}

extern "C" int ether_local_fields(ether_ctx* c, int32_t replica, double* h, double* e) {
    TRY(need_committed(c));
    if (replica < 0 || replica >= c->R) return fail(c, ETHER_ERR_ARG, "ether_local_fields: bad replica");
    const int N = c->N;
    launch_local_fields(meas_args(c), replica, c->d_scratch_h, c->d_scratch_e, c->zee, c->sia,
                        c->model, cst(c));
    HIPCHK(c, hipGetLastError());
    std::vector<float> hh(3 * (size_t)N), ee(N);
    HIPCHK(c, hipMemcpyAsync(hh.data(), c->d_scratch_h, 12 * (size_t)N, hipMemcpyDeviceToHost, cst(c)));
    HIPCHK(c, hipMemcpyAsync(ee.data(), c->d_scratch_e, 4 * (size_t)N, hipMemcpyDeviceToHost, cst(c)));
    HIPCHK(c, hipStreamSynchronize(cst(c)));
    for (int i = 0; i < N; ++i) {
        int s = c->slot_of_site[i];
        if (h)
            for (int x = 0; x < 3; ++x) h[3 * i + x] = hh[3 * s + x];
        if (e) e[i] = ee[s];
    }
    return ETHER_OK;
}

// access for extension translation units (mlp.hip, sq.hip)
namespace ether {
hipStream_t ctx_stream(ether_ctx* c) { return cst(c); }
void*& ctx_mlp(ether_ctx* c) { return c->mlp; }
void*& ctx_sq(ether_ctx* c) { return c->sq; }
int ctx_fail(ether_ctx* c, int code, const std::string& msg) { return fail(c, code, msg); }
int ctx_groups(const ether_ctx* c) { return c->n_groups; }
int ctx_model(ether_ctx* c) { return c->model; }
int ctx_ev_begin(ether_ctx* c, int kind) { return ev_begin(c, kind); }
void ctx_ev_end(ether_ctx* c, int k) { ev_end(c, k); }
int ctx_N(ether_ctx* c) { return c->N; }
int ctx_R(ether_ctx* c) { return c->R; }
int ctx_Rp(ether_ctx* c) { return c->Rp; }
bool ctx_committed(ether_ctx* c) { return c->committed; }
const float* ctx_spin_plane(ether_ctx* c, int comp) {
    return comp == 0 ? c->d_sx : comp == 1 ? c->d_sy : c->d_sz;
}
const int8_t* ctx_ising(ether_ctx* c) { return c->d_s; }
const int* ctx_slot_of_site(ether_ctx* c) { return c->slot_of_site.data(); }
const int* ctx_site_of_slot_dev(ether_ctx* c) { return c->d_site_of_slot; }
int ctx_device(ether_ctx* c) { return c->device; }
bool ctx_dft_resident(ether_ctx* c) { return c->dft_resident != 0; }

// The state a run evolves (every device buffer of the context, the host
// counters of sweeps, swaps and starts, the ladder and its temperatures), to
// put a context back where it was: restoring and running the same schedule
// again gives bit-identical results, since every draw is a function of the
// counters.  Captured graphs are dropped on both sides (they hold counters).
struct CtxSnapshot {
    char* dev = nullptr;
    size_t bytes = 0;
    uint32_t step = 0, init_count = 0, pt_count = 0, dev_step = 0, dev_pt = 0, dev_sc = 0;
    std::vector<double> beta_h;
    int n_groups = 0;
    bool betas_set = false;
};
int ctx_snapshot(ether_ctx* c, void** out) {
    *out = nullptr;
    if (!c->committed || !c->d_arena) return fail(c, ETHER_ERR_STATE, "snapshot of an uncommitted context");
    drop_graphs(c);
    CtxSnapshot* s = new CtxSnapshot;
    hipError_t e = hipMalloc((void**)&s->dev, c->arena_bytes);
    if (e != hipSuccess) {
        delete s;
        return fail(c, ETHER_ERR_HIP, std::string("snapshot: hipMalloc failed: ") + hipGetErrorString(e));
    }
    s->bytes = c->arena_bytes;
    e = hipMemcpyAsync(s->dev, c->d_arena, s->bytes, hipMemcpyDeviceToDevice, cst(c));
    if (e == hipSuccess) e = hipStreamSynchronize(cst(c));
    if (e != hipSuccess) {
        (void)hipFree(s->dev);
        delete s;
        return fail(c, ETHER_ERR_HIP, std::string("snapshot: copy failed: ") + hipGetErrorString(e));
    }
    s->step = c->step;
    s->init_count = c->init_count;
    s->pt_count = c->pt_count;
    s->dev_step = c->dev_step;
    s->dev_pt = c->dev_pt;
    s->dev_sc = c->dev_sc;
    s->beta_h = c->beta_h;
    s->n_groups = c->n_groups;
    s->betas_set = c->betas_set;
    *out = s;
    return ETHER_OK;
}
int ctx_restore(ether_ctx* c, void* snap) {
    const CtxSnapshot* s = static_cast<const CtxSnapshot*>(snap);
    if (!s || s->bytes != c->arena_bytes) return fail(c, ETHER_ERR_STATE, "restore: no snapshot of this context");
    drop_graphs(c);
    hipError_t e = hipMemcpyAsync(c->d_arena, s->dev, s->bytes, hipMemcpyDeviceToDevice, cst(c));
    if (e == hipSuccess) e = hipStreamSynchronize(cst(c));
    if (e != hipSuccess) return fail(c, ETHER_ERR_HIP, std::string("restore: copy failed: ") + hipGetErrorString(e));
    c->step = s->step;
    c->init_count = s->init_count;
    c->pt_count = s->pt_count;
    c->dev_step = s->dev_step;
    c->dev_pt = s->dev_pt;
    c->dev_sc = s->dev_sc;
    c->beta_h = s->beta_h;
    c->n_groups = s->n_groups;
    c->betas_set = s->betas_set;
    return ETHER_OK;
}
void ctx_snapshot_free(void* snap) {
    CtxSnapshot* s = static_cast<CtxSnapshot*>(snap);
    if (!s) return;
    (void)hipFree(s->dev);
    delete s;
}
}  // namespace ether
