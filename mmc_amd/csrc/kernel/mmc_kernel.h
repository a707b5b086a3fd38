/*
 * mmc_kernel.h -- device data layout shared by the transport kernels and the
 * host runtime of the MI355X photon tracer.
 */
#ifndef MMC_KERNEL_H
#define MMC_KERNEL_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mmc_hip.h"
#include "mmc_hip_host.h"

namespace mmc {

constexpr int kBlock = 256;          /* default threads per workgroup (4 waves) */
constexpr int kMaxProp = 4000;       /* media labels 1..kMaxProp per mesh (MAX_PROP, src/mmc_const.h:43) */
constexpr int kPackedProp = 15;      /* labels that fit the 4 low bits of a packed face neighbour */
constexpr int kLaneSlots = 8;        /* per-lane LDS accumulators (energy, Kahan terms) */
constexpr int kMaxReplicas = 64;     /* FP32 field copies: 1..8, or 8 per die up to 64 */
constexpr int kMaxSrc = 8;           /* patterns shared by one photon (srcnum) */

/* Per-simulation parameters read on every step or crossing; staged in LDS
 * (48 B).  flags: bit 0 roulette, 1 energy output, 2 save detectors,
 * 3 voidtime, bits 4-5 boundary condition (MMC_BC_*), bits 6-7 isspecular. */
enum { kFlagRoulette = 1, kFlagEnergy = 2, kFlagSaveDet = 4, kFlagVoidTime = 8 };
constexpr int kReflShift = 4, kSpecShift = 6;
/* bits 8-9: replay output (0 none, 1 Jacobian, 2 weighted path, 3 weighted
 * scattering count); kFlagEnergy marks outputs deposited as raw weight
 * (energy, weighted path, weighted scattering count) rather than fluence */
constexpr int kReplayShift = 8;
enum { kReplayNone = 0, kReplayJ = 1, kReplayWL = 2, kReplayWP = 3 };
/* bit 10: record trajectories; bit 11: momentum transfer block in the
 * detected records */
constexpr int kFlagTraj = 1 << 10;
constexpr int kFlagMom = 1 << 11;
struct SimHot {
    float tstart, tend, rtstep, minenergy;
    float roulettesize;
    int32_t maxgate, flags;
    float nout;                      /* refractive index outside the mesh */
    uint64_t fieldoff;               /* float offset of this simulation's field */
    int32_t srcnum;                  /* field slots per value: patterns sharing each photon */
    int32_t patbase;                 /* this simulation's patterns in KArgs::pattern */
};

/* Per-simulation parameters read at launch / exit / deposit (global, cached). */
/* sets the message mmc_hip_last_error returns; returns code */
int report_error(int code, const char* msg);

struct SimCold {
    float4 srcpos, srcdir, srcparam1, srcparam2, bary0;
    float tstep, nout, minenergy, roulettesize;
    int32_t e0, srctype, detnum, detbase;
    int32_t patbase, patnx, patny, srcnum;
    int32_t isreflect, isspecular, voidtime, outputtype;
    int32_t issavedet, issaveexit, pad0, pad1;
    int64_t replaybase;              /* first of this simulation's photons in KArgs::seeds, or -1 */
    int64_t nphoton, photon_first;
    uint64_t seed;
    uint64_t fieldoff;               /* float offset of this simulation's field */
};

/* Dual-grid output (DMMC) of one simulation: voxel (ix, iy, iz) of edge
 * step at origin org holds value iz*dxy + iy*dx + ix of each gate. */
struct SimGrid {
    float4 org;                      /* grid origin; w = 1/step */
    float step;
    int32_t dx, dxy, ncells;         /* nx, nx*ny, nx*ny*nz */
    int32_t nx, ny, nz, pad;
};

/* Accumulators per simulation (FP64, final reduction only), followed by the
 * launched and absorbed energy of each shared pattern. */
enum SimAcc { kAccLaunch = 0, kAccAbsorb, kAccEscape, kAccRGain, kAccRaytet, kAccCount };
constexpr int kAccSrcLaunch = kAccCount, kAccSrcAbsorb = kAccCount + kMaxSrc;
constexpr int kAccStride = kAccCount + 2 * kMaxSrc;

struct KArgs {
    /* mesh, structure of arrays */
    const float4* planes;            /* ne*4: nx, ny, nz, d */
    const int4* facenb;              /* ne: (neighbour << lblshift) | its label when lblshift = 4 */
    int32_t lblshift;                /* 4: labels packed into facenb; 0: plain neighbour ids, labels
                                        read from type (more than 15 labels, or 2^27+ elements) */
    const int4* elem;                /* ne, 1-based */
    const int32_t* type;             /* ne */
    const float4* node;              /* nn, w unused */
    const float4* enode;             /* ne*3: the 12 node coordinates of each element (Plucker) */
    const int32_t* srcelem;
    int32_t srcelemlen;
    /* launch locator over srcelem: uniform grid, CSR lists of positions in
     * srcelem (ascending), or null for a linear search */
    const int32_t* loc_start;
    const int32_t* loc_elem;
    float4 loc_min, loc_inv;
    int4 loc_dim;
    int32_t nn, ne, prop, nsim;
    int32_t isextdet;                /* wide-field detector elements labelled prop+1 */
    int32_t mom;                     /* 1: lanes keep a momentum-transfer block of prop values */
    /* simulations */
    const SimHot* hot;               /* nsim */
    const SimCold* cold;             /* nsim */
    const float4* media;             /* nsim*(prop+1) */
    const int64_t* cum;              /* nsim+1 cumulative photon counts */
    const float4* detpos;
    const float* pattern;
    /* outputs: nrep copies of the FP32 field, repstride floats apart; a
     * workgroup deposits into copy (XCD id mod nrep) */
    float* field;
    float* field_lo;                 /* compensation buffer (TwoSum atomics) or null */
    uint64_t repstride;
    int32_t nrep, pad_rep;
    double* simacc;                  /* nsim*kAccStride */
    float* detdata;
    int64_t* detphoton;
    unsigned long long* detcount;    /* nsim+2 counters: [0] global slot counter, [1+i] detections of
                                        simulation i, [nsim+1] trajectory records */
    int64_t maxdet;
    int32_t reclen, savedet;
    mmc_hip_photon* rec;
    int64_t recfirst, reccount;      /* global photon index window */
    /* work queue */
    unsigned long long* queue;
    int64_t total;
    int32_t chunk;
    uint32_t maxsteps;               /* per-photon bound on ray-tetrahedron tests */
    int32_t maxsrc;                  /* largest srcnum of the launch (per-lane pattern slots) */
    /* replayed photons: initial xorshift128+ states, detected weights and
     * times, per photon of the launch (SimCold::replaybase) */
    const uint64_t* seeds;
    const float* rweight;
    const float* rtime;
    const SimGrid* grid;             /* nsim, dual-grid tracer only */
    float* traj;                     /* maxtraj trajectory records, or null */
    uint64_t maxtraj;
    float omega;                     /* RF modulation (rad/s), RF kernels only */
    int32_t pad_rf;
};

/* Per-photon xorshift128+ state: splitmix64 finalizer over (seed, photon id).
 * Host and device use the same rule, so a CPU run can replay any photon. */
__host__ __device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
__host__ __device__ __forceinline__ void photon_seed(uint64_t seed, uint64_t id, uint64_t& s0,
                                                     uint64_t& s1) {
    uint64_t z = seed + (2 * id + 1) * 0x9E3779B97F4A7C15ull;
    s0 = mix64(z);
    s1 = mix64(z + 0x9E3779B97F4A7C15ull);
    if ((s0 | s1) == 0) s1 = 1;
}

/* Tracer and output of a kernel variant: branch-less Badouel with element
 * output, Plucker with nodal output, branch-less Badouel with dual-grid
 * (voxel) output, and the two Badouel outputs with RF (complex) weights. */
enum { kTrElem = 0, kTrNodal = 1, kTrGrid = 2, kTrElemRF = 3, kTrGridRF = 4, kTrElemNodal = 5 };
/* where a launch keeps its per-simulation tables: read through the caches
 * or staged in LDS */
enum { kTabGlobal = 0, kTabLds = 1 };

/* OCC: minimum resident workgroups of kBlock threads per CU the register
 * allocation must allow (4 = 128 VGPRs, the default; 5 trades registers for
 * a fifth workgroup).  TAB: per-simulation tables staged in LDS (small
 * batches) or read through the caches (large batches, whose tables would
 * crowd LDS). */
template <int TR, bool COMP, int OCC, int TAB>
__global__ void transport(KArgs a);
__global__ void field_reduce(float* f, float* lo, double* out, uint64_t n, double* sum, int nrep,
                             uint64_t repstride);
__global__ void fill_zero_u64(unsigned long long* p, int n);
/* host C library agreement tables (mmc_host_libm_table) into the current
 * device's copy; 0 or a negative error */
int upload_libm_tables(const uint8_t* blog, const uint8_t* baz, const uint8_t* biso, const uint8_t* bpi,
                       const float* tlog, const float* taz, const float* tiso, const float* tpi);

}  // namespace mmc

#endif
