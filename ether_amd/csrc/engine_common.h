// engine_common.h - device layout and kernel argument blocks of the sweep engine.
//
// Device layout (per context, all replicas resident in HBM):
//   spins   replica-major: element (slot, r) at slot*Rp + r, Rp = R rounded up
//           to 64.  XYZ: three FP32 planes sx, sy, sz (structure of arrays);
//           ANGLE adds an FP32 phi plane.  Ising: one int8 plane of +-1.
//           A wavefront holds one site of 64 consecutive replicas, so every
//           spin load and store of a wave is one contiguous 256-byte (FP32)
//           or 64..256-byte (int8) segment.
//   ell     neighbour table ell[slot*D + k], k < D: low 24 bits = neighbour
//           slot, high 8 bits = coupling class (class 0 = padding, zero
//           coupling).  Shared by all replicas; wave-uniform, so a wave keeps
//           its row in scalar registers and addresses neighbour rows on the
//           scalar unit (buffer loads with a scalar row offset).
//   classes XYZ: FP32 (Jxx,Jyy,Jzz) per class.  Ising: one J per class and a
//           mixed-radix stride so that sum_k stride_k * n_k (n_k = sum of the
//           class-k neighbour spins) indexes an exact threshold table.
//   slots   sites are renumbered so that every colour class is a contiguous
//           slot range; a colour phase writes one range and reads only the
//           others, which is what makes the in-place update race free.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

namespace ether {

constexpr int kBlock = 256;          // threads per workgroup
constexpr int kWave = 64;
constexpr int kWavesPerBlock = kBlock / kWave;
constexpr int kDefaultSitesPerWave = 4;  // consecutive sites of a colour class per wave
constexpr int kMeasRecordSpw = 4;       // least site steps per wave of a sweep writing measurement records
constexpr int kMaxSitesPerWave = 16;     // (a launch parameter: ether_set_launch / ether_autotune)
constexpr int kMaxClasses = 255;     // coupling classes (8-bit field of an ELL entry)
constexpr int kMaxSpecies = 16;
constexpr int kMaxIsingClasses = 8;
constexpr int kIsingLdsTable = 4096; // thresholds staged in LDS up to this many entries
constexpr int kMeasWords = 6;        // fused measurement record: E_exch, E_z, E_sia, Mx, My, Mz
constexpr int kMeasCopies = 16;      // copies of the fused sums the workgroups spread their atomics over
constexpr int kAccCopies = 16;       // copies of the acceptance counters, likewise
static_assert(kMeasCopies <= kWave, "the finalize reads one copy per lane");

struct XyzArgs {
    float* sx;
    float* sy;
    float* sz;
    float* phi;
    const uint32_t* ell;   // [N][D]
    int D;
    const uint8_t* sp;
    int N;
    int R;
    int Rp;
    int G;                 // replica groups of 64 = Rp/64
    int c_begin;
    int c_end;
    int nblocks;
    int nchunk;            // site chunks of the phase: nblocks = groups x nchunk (+ padding)
    int xcd_map;
    int spw;               // steps per wave (sites per wave = spw * sps)
    int lps;               // lanes per site (64: one site per wave); lane l of a wave holds
                           // replica l % lps of site l / lps, lanes from sps * lps idle
    int sps;               // sites per step of a wave = 64 / lps
    const int* slot_of;
    const float* beta_slot;
    uint64_t seed;
    const uint32_t* ctr;   // device counters: [0] sweep base, [1] swap base, [2] swap count
    uint32_t step_off;     // sweep counter = ctr[0] + step_off
    uint32_t pass;
    float frac;
    int fuse;              // sweep: reflect (with probability frac) before the proposal
    uint32_t plane_bytes;  // bytes of one spin plane, N * Rp * 4 (< 2 GiB)
    unsigned long long* acc;   // [kAccCopies][R slots] accepted moves
    int n_classes;
    const float* jclass;
    int nsp;
    const float* zvec;
    const float* dsia;
    float dphi;
    int zee;               // Zeeman term on (wave-uniform runtime flag)
    int sia;               // single-ion anisotropy on
    // Fused measurement (set on the phases of a sweep that a measurement
    // follows).  Each updated site adds the terms of its final spin to the
    // record {E_exch, E_z, E_sia, Mx, My, Mz} of its replica: where all its
    // neighbours are of earlier colours (final already) the exchange energy
    // of its bonds, counted twice; its Zeeman and SIA energies and its
    // weighted spin.  A term is rounded per site to an integer multiple of
    // 1/mscale and summed exactly (int64 in macc), so the totals do not
    // depend on which wave or workgroup adds what, nor in which order.
    //
    // With three or more colours a phase's sites have neighbours of earlier
    // and of later colours: a site books its bonds to the earlier ones (the
    // exchange field of those alone, from the slots below c_begin), each
    // bond again once, from its later-coloured end.  Each workgroup then
    // writes its sums to a record of its own (mrec) instead of adding them to
    // macc, and the finalize adds the records up.
    int meas;              // 0: off
    int meas_lower;        // 0: no neighbour of an earlier colour, 1: all of them, 2: some
    const float* w;        // [N] magnetisation weight per slot
    unsigned long long* macc;  // [kMeasCopies][Rp][kMeasWords] two's-complement int64 sums
    long long* mrec;       // non-null: records [phase][group][lane][mrec_pp units][kMeasWords]
    int mrec_stride;       // workgroups of every phase of the measuring sweep
    int mrec_pp;           // unit slots per (phase, group, lane): unit bb is slot bb / G of group bb % G
    int mrec_phase;
    double mscale[4];      // powers of two for exchange, Zeeman, SIA, magnetisation
};

struct IsingArgs {
    int8_t* s;             // [N][Rp]
    const uint32_t* ell;   // [N][D]
    int D;
    const uint8_t* sp;
    int N;
    int R;
    int Rp;
    int G4;                // groups of 256 replicas (4 per lane)
    int c_begin;
    int c_end;
    int nblocks;
    int nchunk;            // site chunks of the phase: nblocks = groups x nchunk (+ padding)
    int xcd_map;
    int spw;               // steps per wave (sites per wave = spw << (6 - lps_shift))
    int lps_shift;         // lanes per site = 1 << lps_shift, 4 replicas per lane (6: one site)
    const int* slot_of;
    uint64_t seed;
    const uint32_t* ctr;
    uint32_t step_off;
    unsigned long long* acc;   // [kAccCopies][R slots] accepted moves
    int n_classes;
    const int* stride;     // [n_classes+1], stride[0] = 0
    int offset;            // sum_k Z_k * stride_k
    int table_per_slot;    // nsp * 2 * P
    int P;
    int table_in_lds;      // the whole [R][table] fits kIsingLdsTable
    const uint32_t* thr;   // [R slots][nsp][2][P]
};

// Ising sweeps with a replica word's whole lattice resident in LDS: one
// workgroup per group of 4 replicas (one 32-bit word of int8 spins per
// site), n_sweeps sweeps of every colour phase with workgroup barriers
// between the phases, and optionally the measurement of the final state.
constexpr int kMaxResColours = 16;
constexpr size_t kResidentLds = 60 * 1024;   // LDS per workgroup of the resident kernel (dynamic part)
struct IsingResArgs {
    int8_t* s;             // [N][Rp]
    const uint32_t* ell;   // [N][D]
    int D;
    const uint8_t* sp;
    const float* w;        // [N] magnetisation weight per slot
    int N;
    int R;
    int Rp;
    int ncol;
    int cbeg[kMaxResColours + 1];
    const int* slot_of;
    uint64_t seed;
    const uint32_t* ctr;
    uint32_t step_off;     // sweep k of the launch draws with step ctr[0] + step_off + k
    int n_sweeps;
    unsigned long long* acc;
    int n_classes;
    int stride[kMaxIsingClasses + 1];
    int offset;
    int table_per_slot;
    int P;
    int table_in_lds;      // the four replicas' tables are staged in LDS
    const uint32_t* thr;
    int meas;              // book the final state's measurement record (tile 0 of partial)
    int nsp;
    int rec;
    double* partial;
};

struct MeasArgs {
    const float* sx;
    const float* sy;
    const float* sz;
    const int8_t* s;
    const uint32_t* ell;
    int D;
    const uint8_t* sp;
    const float* w;
    int N;
    int R;
    int Rp;
    int G;
    int tile_sites;        // sites per wave unit
    int tiles;
    int rec;               // doubles per partial record
    int n_classes;
    const float* jclass;
    int nsp;
    const float* zvec;
    const float* dsia;
    // measurement rows: every bond once, from its endpoint with the smaller
    // site index (bond_factor 2), or the full rows where a site bonds to
    // itself (bond_factor 1); padding entries are 0 (slot 0, class 0, zero
    // coupling)
    const uint32_t* ell_m;
    const uint32_t* ell_m_n;   // [N] entries used in each measurement row
    int D_m;
    int bond_factor;
    uint32_t plane_bytes;
    const int* spiral;     // [N slots][spiral_cap] partner slots, -1 ends a row (XYZ)
    int spiral_cap;        // 0: no spiral-state observable
    double* partial;       // [tiles][Rp][rec]
};

struct FinArgs {
    const double* partial;
    int tiles;
    int rec;
    int R;
    int Rp;
    int N;
    int model;
    int n_classes;
    const double* ising_j;   // [n_classes+1]
    int nsp;
    const double* ising_zc;  // [nsp]  -g*muB*s*Hz
    int sample;
    int pt_etot;
    double* E_pt;            // [R]
    double* parts;           // [R][4]
    double* M;               // [R][3]
    long long* bonds;        // [R][n_classes+1]
    double* moments;         // [R slots][8]
    const int* slot_of;      // [R]
    double spiral_count;     // > 0: record words 6..8 hold sum S_i x S_j over the partners
    double* psi;             // [R] spiral-state parameter of the measured state
    double* mom_sp;          // [R slots][3] sums of psi, psi^2, psi^4 over samples
    unsigned long long* macc;  // non-null: the record comes from the fused measurement's
                               // [kMeasCopies][Rp][kMeasWords] sums (cleared here for the next one)
    const long long* mrec;     // non-null: it comes from the per-workgroup records of the
    int mrec_stride;           // measuring sweep's mrec_phases phases (XyzArgs::mrec): units
    int mrec_phases;           // bb < mrec_stride, unit bb holding replicas (bb % mrec_groups) * 64 + lane
    int mrec_groups;
    int mrec_pp;
    double mscale[4];
};

struct PtArgs {
    int n_groups;
    const int* gstart;
    int parity;            // 0/1, or -1: (ctr[2] + pt_off + 1) & 1 (the reference's swap_count parity)
    int* rep_at;
    int* slot_of;
    const double* beta_slot;
    const double* E_pt;
    uint64_t seed;
    const uint32_t* ctr;
    uint32_t pt_off;       // swap index = ctr[1] + pt_off
    unsigned long long* acc;
    unsigned long long* att;
    unsigned long long* trips;
    int* trip_state;
    int* n_accepted;
};

struct InitArgs {
    float* sx;
    float* sy;
    float* sz;
    float* phi;
    int8_t* s;
    int N;
    int R;
    int Rp;
    int model;
    int angle;
    int mode;  // 1 fresh, 2 random
    float dphi;
    uint64_t seed;
    uint32_t step;
};

// Workgroup -> wave unit bb, which the kernels split as (replica group
// bb % groups, site chunk bb / groups).  With xcd_map the launch has a
// multiple of 8 workgroups; the dispatcher deals consecutive workgroup ids
// round-robin to the 8 XCDs, so workgroup b runs on XCD b%8.  Map 1 remaps
// b -> p = (b%8)*per + b/8, which gives every XCD one contiguous range of
// units: a range of sites with every replica group.  Map 2 also orders the
// units group-major (p -> group p / nchunk, chunk p % nchunk), so that an XCD
// walks more sites of fewer groups: a group's neighbour halo is 1/groups of
// the spins of those sites, and it stays in the XCD's L2 until the sites
// next to it reuse it.  Units p >= groups * nchunk (padding) keep bb = p,
// whose chunk lies past the colour's sites.
__device__ __forceinline__ int block_index(int nblocks, int xcd_map, int groups, int nchunk) {
    int b = blockIdx.x;
    if (!xcd_map) return b;
    const int per = nblocks >> 3;
    b = (b & 7) * per + (b >> 3);
    if (xcd_map == 2 && b < groups * nchunk) b = (b % nchunk) * groups + b / nchunk;
    return b;
}

inline int launch_blocks(long long units, int xcd_map) {
    long long nb = (units + kWavesPerBlock - 1) / kWavesPerBlock;
    if (xcd_map) nb = (nb + 7) / 8 * 8;
    return (int)nb;
}

// ELL row width: D rounded up to an unrolled bucket (4,6,8,12,18,24,32), or D
// itself above 32 (runtime loop).  Rows are padded with class-0 entries.
int ell_bucket(int D);

// ---------------------------------------------------------------------------
// Launch recording for batched execution.  While g_recorder is set (by the
// batch runner, on the calling host thread), the launchers below append the
// launch they would issue, with its argument block, instead of issuing it.
// The batch runner then issues one launch per (kind, variant) covering the
// recorded launches of every lattice (k_*_batch kernels).
// ---------------------------------------------------------------------------
enum LaunchKind : int {
    LK_XYZ_SWEEP = 0,
    LK_XYZ_OVRR = 1,
    LK_ISING_SWEEP = 2,
    LK_MEASURE = 3,
    LK_FINALIZE = 4,
    LK_PT = 5,
    LK_ISING_RESIDENT = 6,
};
struct LaunchRecord {
    int kind;
    int variant;   // template instance: sweep (meas << 18) | (lanes << 17) | (angle << 16) | D;
                   // measure (wide Ising << 3) | (model << 2) | (zee << 1) | sia
    int units;     // workgroups of the launch on its own, without XCD padding (0: nothing to do)
    int bytes;     // size of the argument block
    int block;     // threads per workgroup (0: kBlock)
    int lds;       // dynamic LDS bytes per workgroup
    union {
        XyzArgs x;
        IsingArgs i;
        IsingResArgs r;
        MeasArgs m;
        FinArgs f;
        PtArgs p;
    } u;
};
struct LaunchRecorder {
    std::vector<LaunchRecord> recs;
};
extern thread_local LaunchRecorder* g_recorder;
// one batched launch: args[n_args] of the record kind, items[n_items] =
// (argument index or -1, local workgroup); block and lds are the largest of
// the recorded launches (LK_ISING_RESIDENT; the other kinds use kBlock and
// static LDS)
void launch_batch(int kind, int variant, const void* d_args, const int2* d_items, int n_items,
                  int block, int lds, hipStream_t st);
// opts a batched kernel instance into `lds` bytes of dynamic LDS (before a
// graph capture records its launch)
void grant_batch_lds(int kind, int variant, int lds);
// dynamic LDS of the batched resident Ising kernel, per workgroup (one per CU)
constexpr size_t kBatchResidentLds = 150 * 1024;
// counters of several contexts: set (vals[3*i..]) or advance by (dstep, dswaps)
void launch_ctr_batch(uint32_t* const* d_ctrs, int n, const uint32_t* d_vals, uint32_t dstep,
                      uint32_t dswaps, hipStream_t st);

// launchers (kernels.hip)
void launch_xyz_sweep(const XyzArgs& a, bool angle, bool zee, bool sia, hipStream_t st);
void launch_xyz_ovrr(const XyzArgs& a, bool angle, bool zee, hipStream_t st);
void launch_ising_sweep(const IsingArgs& a, hipStream_t st);
void launch_ising_resident(const IsingResArgs& a, int block, size_t lds_bytes, bool per_replica, hipStream_t st);
void launch_measure(const MeasArgs& a, int model, bool zee, bool sia, hipStream_t st);
void launch_finalize(const FinArgs& a, hipStream_t st);
void launch_pt_swap(const PtArgs& a, hipStream_t st);
void launch_init(const InitArgs& a, hipStream_t st);
void launch_local_fields(const MeasArgs& a, int r, float* h, float* e, int zee, int sia,
                         int model, hipStream_t st);
void launch_extract(const float* plane, const int8_t* s, int N, int Rp, int r, float* out,
                    hipStream_t st);
// device counters: set to absolute values / advance by deltas (single-thread kernels,
// used instead of memset nodes so that captured graphs replay correctly)
void launch_set_ctr(uint32_t* ctr, uint32_t step, uint32_t swaps, uint32_t swap_count,
                    hipStream_t st);
void launch_advance_ctr(uint32_t* ctr, uint32_t dstep, uint32_t dswaps, hipStream_t st);
void launch_insert(float* plane, int8_t* s, int N, int Rp, int r0, int r1, const float* in,
                   hipStream_t st);

}  // namespace ether
