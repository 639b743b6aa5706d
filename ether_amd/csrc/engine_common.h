// engine_common.h - device layout and kernel argument blocks of the sweep engine.
//
// Device layout (per context, all replicas resident in HBM):
//   spins   XYZ: three FP32 planes sx, sy, sz of R*N floats (structure of
//           arrays, replica-major, colour-sorted slots); ANGLE adds an FP32
//           phi plane.  Ising: one int8 plane of R*N values +-1.
//   ell     neighbour table in ELL form, ell[k*ell_stride + slot] for
//           k < D: low 24 bits = neighbour slot, high 8 bits = coupling
//           class (class 0 = padding, zero coupling).  Shared by replicas.
//   classes XYZ: FP32 (Jxx,Jyy,Jzz) per class.  Ising: one J per class and
//           a mixed-radix stride so that sum_j stride[class_j]*sigma_j
//           indexes an exact acceptance-threshold table.
//   slots   sites are renumbered so that every colour class is a contiguous
//           slot range; a colour phase touches one range and reads only the
//           other ranges, which is what makes the in-place update race free.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ether {

constexpr int kBlock = 256;          // threads per workgroup for sweep/OVRR kernels
constexpr int kMeasPerThread = 4;    // sites per thread in the measurement pass
constexpr int kMeasTile = kBlock * kMeasPerThread;
constexpr int kMaxClasses = 255;     // coupling classes (8-bit field of an ELL entry)
constexpr int kMaxSpecies = 16;
constexpr int kMaxIsingClasses = 8;
constexpr int kIsingLdsTable = 4096; // thresholds staged in LDS up to this many entries
constexpr int kRecord = 32;          // doubles per measurement partial record

struct XyzArgs {
    float* sx;
    float* sy;
    float* sz;
    float* phi;
    const uint32_t* ell;
    int ell_stride;
    int D;
    const uint8_t* sp;
    int N;
    int R;
    int c_begin;
    int c_end;
    int tiles;
    int xcd_map;
    const int* slot_of;
    const float* beta_slot;
    uint64_t seed;
    uint32_t step;
    uint32_t pass;
    float frac;
    unsigned long long* acc;
    int n_classes;
    const float* jclass;
    int nsp;
    const float* zvec;
    const float* dsia;
    float dphi;
};

struct IsingArgs {
    int8_t* s;
    const uint32_t* ell;
    int ell_stride;
    int D;
    const uint8_t* sp;
    int N;
    int R;
    int c_begin;
    int c_end;
    int tiles;
    int xcd_map;
    const int* slot_of;
    uint64_t seed;
    uint32_t step;
    unsigned long long* acc;
    int n_classes;
    const int* stride;       // [n_classes+1], stride[0] = 0
    int offset;              // sum_k Z_k * stride_k
    int table_per_slot;      // nsp * 2 * P
    int P;
    const uint32_t* thr;     // [R slots][nsp][2][P]
};

struct MeasArgs {
    const float* sx;
    const float* sy;
    const float* sz;
    const int8_t* s;
    const uint32_t* ell;
    int ell_stride;
    int D;
    const uint8_t* sp;
    const float* w;
    int N;
    int R;
    int tiles;
    int n_classes;
    const float* jclass;
    int nsp;
    const float* zvec;
    const float* dsia;
    double* partial;         // [R][tiles][kRecord]
};

struct FinArgs {
    const double* partial;
    int tiles;
    int R;
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
};

struct PtArgs {
    int n_groups;
    const int* gstart;
    int parity;
    int* rep_at;
    int* slot_of;
    const double* beta_slot;
    const double* E_pt;
    uint64_t seed;
    uint32_t swap_index;
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
    int model;
    int angle;
    int mode;  // 1 fresh, 2 random
    float dphi;
    uint64_t seed;
    uint32_t step;
};

// blockIdx -> (replica, tile).  With xcd_map, consecutive workgroup ids
// (which the dispatcher deals round-robin to the 8 XCDs) are arranged so that
// every tile of replica r lands on XCD r%8 and shares that XCD's L2.
__device__ __forceinline__ bool map_block(int R, int tiles, int xcd_map, int& r, int& tile) {
    int b = blockIdx.x;
    if (xcd_map) {
        int xcd = b & 7;
        int j = b >> 3;
        r = (j / tiles) * 8 + xcd;
        tile = j % tiles;
    } else {
        r = b / tiles;
        tile = b % tiles;
    }
    return r < R;
}

inline int blocks_for(int R, int tiles, int xcd_map) {
    if (xcd_map) return 8 * ((R + 7) / 8) * tiles;
    return R * tiles;
}

// launchers (kernels.hip)
void launch_xyz_sweep(const XyzArgs& a, bool angle, bool zee, bool sia, hipStream_t st);
void launch_xyz_ovrr(const XyzArgs& a, bool angle, bool zee, hipStream_t st);
void launch_ising_sweep(const IsingArgs& a, hipStream_t st);
void launch_measure(const MeasArgs& a, int model, bool zee, bool sia, hipStream_t st);
void launch_finalize(const FinArgs& a, hipStream_t st);
void launch_pt_swap(const PtArgs& a, hipStream_t st);
void launch_init(const InitArgs& a, hipStream_t st);
void launch_local_fields(const MeasArgs& a, int r, float* h, float* e, int zee, int sia,
                         int model, hipStream_t st);

}  // namespace ether
