/*
 * mmc_hip.h -- C ABI of the MI355X photon-transport library (libmmc_hip.so).
 *
 * The reference selects a transport backend with one call of the shape
 *     void mmc_run_XX(mcconfig* cfg, tetmesh* mesh, raytracer* tracer)
 * (mmc_run_cl: src/mmc_cl_host.h:126, mmc_run_cu: src/mmc_cu_host.h:58,
 *  mmc_run_mp: src/mmc_host.h:54), dispatched in src/mmc.c:82-103.
 * This header exposes the same operation with plain pointers and sizes: a
 * caller fills mmc_hip_mesh from tetmesh/raytracer after mmc_prep()
 * (src/mmc_host.c:136-165) and mmc_hip_sim from mcconfig, and receives the
 * fluence field, detected-photon records and energy totals that mmc_run_cl
 * writes into cfg->exportfield / exportdetected / energytot / energyesc.
 *
 * All element and node ids are 1-based, as in tetmesh.  Nothing in this ABI
 * refers to torch or to HIP types.
 */
#ifndef MMC_HIP_H
#define MMC_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MMC_HIP_ABI_VERSION 8

/* ray-tracing methods (enum TRTMethod, src/mmc_utils.h:97) */
#define MMC_RT_PLUCKER   0
#define MMC_RT_BLBADOUEL 3
/* dual-grid MMC (DMMC, "-M G", the reference's GPU default): branch-less
 * Badouel tracing with the deposits on a voxel grid over the mesh's bounding
 * box (src/mmc_raytrace.c:1615-1673, src/mmc_mesh.c:349-381) */
#define MMC_RT_BLBADOUELGRID 4
/* source types (enum TSrcType, src/mmc_utils.h:101) */
#define MMC_SRC_PENCIL      0
#define MMC_SRC_ISOTROPIC   1
#define MMC_SRC_CONE        2
#define MMC_SRC_GAUSSIAN    3
#define MMC_SRC_PLANAR      4
#define MMC_SRC_PATTERN     5
#define MMC_SRC_FOURIER     6
#define MMC_SRC_ARCSINE     7
#define MMC_SRC_DISK        8
#define MMC_SRC_FOURIERX    9
#define MMC_SRC_FOURIERX2D  10
#define MMC_SRC_ZGAUSSIAN   11
#define MMC_SRC_LINE        12
#define MMC_SRC_SLIT        13
#define MMC_SRC_PENCILARRAY 14
#define MMC_SRC_PATTERN3D   15
#define MMC_SRC_HYPERBOLOID 16
#define MMC_SRC_RING        17
/* output types (enum TOutputType, src/mmc_utils.h:106) */
#define MMC_OUT_FLUX     0
#define MMC_OUT_FLUENCE  1
#define MMC_OUT_ENERGY   2
/* replay outputs (enum TOutputType, src/mmc_utils.h:106): photons relaunched
 * from saved seeds deposit the detected-weight-scaled Jacobian for mua (J),
 * weighted path length (L) or weighted scattering count (P) */
#define MMC_OUT_JACOBIAN 3
#define MMC_OUT_WL       4
#define MMC_OUT_WP       5
/* debug flag that records photon trajectories (dlTraj, "-D M",
 * src/mmc_utils.h:92-94) */
#define MMC_DL_TRAJ      8192
/* floats per trajectory record: photon id (uint32 bits), x, y, z, weight,
 * element id (int32 bits) (MCX_DEBUG_REC_LEN, src/mmc_const.h:47) */
#define MMC_TRAJ_RECLEN  6
/* boundary conditions (enum TBoundary, src/mmc_utils.h:116) */
#define MMC_BC_NOREFLECT 0
#define MMC_BC_REFLECT   1
#define MMC_BC_ABSORBEXT 2
#define MMC_BC_MIRROR    3

/* per-photon termination codes in mmc_hip_photon.reason */
#define MMC_END_EXIT     1   /* left the mesh through an exterior face */
#define MMC_END_LOST     2   /* no exit face after 3 nudges (eid = 0x7FFFFFFF) */
#define MMC_END_ROULETTE 3   /* killed by Russian roulette */
#define MMC_END_STEPLIMIT 4  /* exceeded the per-photon step bound (a photon cycling at a vertex) */
#define MMC_MAX_STEPS    (1u << 26)  /* per-photon bound on ray-tetrahedron tests */

/* Mesh after mmc_prep: the fields of tetmesh (src/mmc_mesh.h:87-122) and
 * raytracer->n (src/mmc_mesh.h:135-141) that the transport reads. */
typedef struct mmc_hip_mesh {
    int32_t nn, ne, prop, nf;
    const float* node;       /* nn*3 */
    const int32_t* elem;     /* ne*4, positively oriented (mesh_getvolume) */
    const int32_t* type;     /* ne, 0 = background, 1..prop */
    const int32_t* facenb;   /* ne*4, exterior faces numbered -1,-2,... (tracer_prep) */
    const float* planes;     /* ne*16 per element {nx[4],ny[4],nz[4],d[4]} (tracer->n) */
    const float* med;        /* (prop+1)*4 {mua,mus,g,n}; row 0 = background (mesh->med) */
    const int32_t* srcelem;  /* candidate launch elements for area sources (mesh->srcelem) */
    int32_t srcelemlen;
    /* 1: elements labelled prop+1 are wide-field detector elements (mesh label -2,
     * mesh_srcdetelem, src/mmc_mesh.c:390-427); they take the background's optical
     * properties, photons leaving the tissue keep travelling through the background,
     * and one that leaves the mesh through a detector element is detected with that
     * element's id as detector id (point detectors are then ignored) */
    int32_t isextdet;
} mmc_hip_mesh;

/* One simulation: the mcconfig fields (src/mmc_utils.h:210-345) the
 * transport path reads, plus the per-photon seed of this build. */
typedef struct mmc_hip_sim {
    int64_t nphoton;
    uint64_t seed;           /* photon i starts from mmc_hip_photon_seeds(seed, i) */
    int64_t photon_offset;   /* id of the first photon (respin / chunked runs) */
    int32_t method;          /* MMC_RT_PLUCKER, MMC_RT_BLBADOUEL or MMC_RT_BLBADOUELGRID */
    int32_t basisorder;      /* 0 element field, 1 nodal field (ignored by the dual grid) */
    int32_t outputtype;
    int32_t isreflect, isspecular, voidtime;
    int32_t srctype, srcnum;
    int32_t e0;              /* initial element (cfg->e0) */
    float bary0[4];          /* its barycentric coordinates (cfg->bary0) */
    float srcpos[4], srcdir[4], srcparam1[4], srcparam2[4];
    const float* srcpattern; /* pattern source: srcparam1.w*srcparam2.w pixels x srcnum patterns,
                                pattern k of pixel p at [p*srcnum + k]; with srcnum > 1 each
                                photon is shared by all patterns and the field holds srcnum
                                slots per value, [gate][node or element][pattern] */
    float tstart, tend, tstep;
    int32_t maxgate;
    float nout, minenergy, roulettesize;
    int32_t issavedet, issaveexit;
    int32_t detnum;
    const float* detpos;     /* detnum*4 {x,y,z,r} */
    int64_t maxdetphoton;
    const float* med;        /* optional (prop+1)*4 media replacing mesh->med */
    /* replay (seed = SEED_FROM_FILE, src/mmc_mesh.c:815-898): photon i of the
     * simulation starts from photonseed[2i], photonseed[2i+1] (xorshift128+
     * state; photon_offset selects the first) instead of the seeded hash;
     * replayweight/replaytime give each photon's detected weight and time of
     * flight for outputs J/L/P.  NULL otherwise. */
    const uint64_t* photonseed;
    const float* replayweight;
    const float* replaytime;
    /* dual grid: voxel edge in mesh units (cfg->steps.x, "--gridsize",
     * Domain.Step); the grid covers the mesh's bounding box widened by 1e-6
     * (mmc_hip_sim_grid / mmc_host_grid give its origin and dimensions) */
    float gridstep;
    float unitinmm;          /* mesh length unit in mm (normalization of grid fields); 0 = 1 */
    int32_t debuglevel;      /* MMC_DL_TRAJ: record each photon's launch, scattering and end
                                positions (savedebugdata, src/mmc_raytrace.c:1739-1756) */
    /* RF forward (cfg->omega, rad/s; mmc_core.cl IS_RF): with omega > 0 the photon carries a
     * complex weight turning by omega*n/c0 per unit path and deposits the complex fluence;
     * the imaginary part goes to result.field_im.  Branch-less Badouel (element or grid
     * output) with srcnum 1; ignored for replayed photons, as in the reference.  One
     * omega per batch. */
    float omega;
    /* cfg->ismomentum (-m): detected records carry, after the partial paths, the momentum
     * transfer sum of (1 - cos theta) over the photon's scattering events per medium
     * (src/mmc_mesh.c:1707-1709, src/mmc_raytrace.c:2073-2075) */
    int32_t ismomentum;
} mmc_hip_sim;

/* Per-photon end state, written for photon ids in [recfirst, recfirst+reccount). */
typedef struct mmc_hip_photon {
    float p[3]; float w;     /* final position and weight (0 after roulette death) */
    float v[3]; float t;     /* final direction, time of flight (s) */
    int32_t eid, reason, nraytet, detid;
    float eabs, rgain, w0, nscat;  /* absorbed energy, roulette gain, launch weight, scatterings */
} mmc_hip_photon;

typedef struct mmc_hip_result {
    double* field;           /* datalen*maxgate, datalen = nn (basisorder 1), ne, or the voxels of
                                the dual grid (nx*ny*nz, x fastest); accumulated (+=);
                                NULL keeps the field on the device (fieldsum is still filled) */
    double fieldsum;         /* sum of the raw (unnormalized) field of this call, FP64 */
    double launchweight;     /* total launched weight   (cfg->energytot) */
    double absorbweight;     /* total absorbed weight */
    double escapeweight;     /* final weight of photons that ended by exit or lost */
    double roulettegain;     /* net weight created (>0) or removed (<0) by roulette */
    double raytet;           /* ray-tetrahedron tests (reporter->raytet) */
    float* detdata;          /* maxdetphoton*reclen records, or NULL */
    int64_t* detphoton;      /* photon id of each saved record, or NULL */
    int64_t detected;        /* detected photons (may exceed maxdetphoton) */
    int32_t reclen;          /* floats per record: 1 + 2*prop + 6*issaveexit + 1 */
    mmc_hip_photon* photons; /* reccount records, or NULL */
    int64_t recfirst, reccount;
    double kernel_ms;        /* transport kernels only */
    double total_ms;         /* including uploads, reductions and copies */
    double* srcweight;       /* pattern sources with srcnum > 1: launched energy of each pattern,
                                then absorbed energy of each pattern (2*srcnum, +=), or NULL */
    float* traj;             /* maxtraj records of MMC_TRAJ_RECLEN floats (cfg->exportdebugdata), or
                                NULL; one buffer per launch, filled through the first result of a
                                batch that provides one, in device order (each photon's records in
                                order) */
    int64_t maxtraj;         /* capacity in records (cfg->maxjumpdebug) */
    int64_t ntraj;           /* records produced (may exceed maxtraj: the rest were dropped) */
    double* field_im;        /* RF: imaginary part of the field, same layout as field (+=), or NULL
                                (cfg->exportadjoint of the reference's RF forward mode) */
} mmc_hip_result;

/* Launch shape for the persistent kernel; zeros pick the defaults. */
typedef struct mmc_hip_tune {
    int32_t block;           /* threads per workgroup (multiple of 64) */
    int32_t blocks_per_cu;   /* resident workgroups per CU */
    int32_t chunk;           /* photons claimed per lane refill round (1 = per photon) */
    int32_t compensated;     /* 1: TwoSum-compensated FP32 atomics (hi+lo buffers) */
    int32_t replicas;        /* FP32 field copies: 1..8 (one per XCD group) or a multiple of 8
                                up to 64 (several per XCD); 0 = automatic: up to 64 while
                                they take at most 256 MiB, else 8 when they fit the
                                device-memory budget, else 4, 2, 1 */
} mmc_hip_tune;

typedef struct mmc_hip_meshres* mmc_hip_mesh_handle;

/* ---- library / device ---------------------------------------------------- */
int32_t mmc_hip_abi_version(void);
const char* mmc_hip_last_error(void);
/* Fills name (len bytes) and returns the CU count, or -1 without a device. */
int32_t mmc_hip_device_info(char* name, int32_t len, int64_t* hbm_bytes);

/* Per-photon xorshift128+ states (2 words per photon) for photons
 * [first, first+n) of a simulation seeded with `seed`.  Host only. */
void mmc_hip_photon_seeds(uint64_t seed, int64_t first, int64_t n, uint64_t* out);

/* ---- plain drop-in (replaces mmc_run_cl, src/mmc_cl_host.c:87-1721) ------- */
/* Uploads the mesh, runs the simulation, accumulates into res, frees all
 * device memory.  Returns 0 or a negative error (see mmc_hip_last_error). */
int32_t mmc_hip_run(const mmc_hip_mesh* mesh, const mmc_hip_sim* sim,
                    const mmc_hip_tune* tune, mmc_hip_result* res);

/* ---- resident meshes and batching (no reference counterpart; the batch
 *      follows the per-slot output offsets of src/mmc_core.cl:859-861) ------ */
int32_t mmc_hip_mesh_upload(const mmc_hip_mesh* mesh, mmc_hip_mesh_handle* out);
int32_t mmc_hip_mesh_release(mmc_hip_mesh_handle h);
int64_t mmc_hip_mesh_bytes(mmc_hip_mesh_handle h);
/* Runs nsim simulations on one resident mesh in a single persistent launch. */
int32_t mmc_hip_run_batch(mmc_hip_mesh_handle h, const mmc_hip_sim* sims, int32_t nsim,
                          const mmc_hip_tune* tune, mmc_hip_result* res);
/* Device bytes the outputs of a simulation occupy (admission accounting). */
int64_t mmc_hip_sim_output_bytes(mmc_hip_mesh_handle h, const mmc_hip_sim* sim);
/* Adjoint Jacobians on the dual grid (the -O A / D / W outputs; mmc_adjoint_kernel and
 * mmc_adjoint_dcoeff_kernel, src/mmc_core.cl:2228-2400, scaled as src/mmc_cl_host.c:1436-1461).
 * field: ns source fields then nd detector fields (each detector run as a source), laid out
 * [slot][gate][voxel] with dim[0]*dim[1]*dim[2] voxels, x fastest (normalized grid fields, as
 * float); field_im: their imaginary parts (RF) or NULL.  jmua / jd (either may be NULL) receive
 * [s][d][voxel], followed by the imaginary block for RF:
 *     J_mua = -unitinmm^3 * cw_s * cw_d,  J_D = -unitinmm * grad(cw_s) . grad(cw_d),
 * cw = the field summed over gates, gradients by second-order finite differences.  0 on success. */
int32_t mmc_hip_adjoint_grid(const float* field, const float* field_im, const int32_t dim[3], int32_t maxgate,
                             int32_t ns, int32_t nd, float unitinmm, float* jmua, float* jd);
/* Adjoint Jacobians on the tetrahedral mesh from nodal fields (basisorder 1;
 * mmc_adjoint_mesh_full_kernel / mmc_adjoint_mesh_nodal_kernel, src/mmc_core.cl:2403-2649, host
 * src/mmc_cl_host.c:1117-1320).  field / field_im: [slot][gate][node] (ns sources, then nd
 * detectors); elem: 1-based node ids, 4 per element; evol: element volumes; deldotdel: per
 * element the packed <grad phi_i . grad phi_j> * evol of mmc_host_deldotdel, as float (needed for
 * jd); nvol: node volumes (needed for mode 1).  mode 0: linear-basis integrals per element,
 * a quarter to each of its nodes (the reference's default); mode 1: J_mua = -nvol * cw_s * cw_d
 * per node (the reference's adjointmode 1; J_D as in mode 0); mode 2: the per-element values of
 * mode 0 ([s][d][elem]).  jmua / jd (either may be NULL) receive [s][d][node] (or elem), then the
 * imaginary block for RF.  0 on success, -1 for bad arguments. */
int32_t mmc_hip_adjoint_mesh(const float* field, const float* field_im, int32_t nn, int32_t ne,
                             const int32_t* elem, const float* evol, const float* deldotdel, const float* nvol,
                             int32_t maxgate, int32_t ns, int32_t nd, int32_t mode, float* jmua, float* jd);
/* Origin and voxel counts of a dual-grid simulation's output on a resident
 * mesh (the grid cfg->dim / mesh->nmin of mesh_createdualmesh,
 * src/mmc_mesh.c:349-381); 0 on success. */
int32_t mmc_hip_sim_grid(mmc_hip_mesh_handle h, const mmc_hip_sim* sim, float nmin[3], int32_t dim[3]);

/* ---- continuous batching (no reference counterpart): simulations are run in
 *      photon chunks whose FP32 fields stay in a caller-managed device arena
 *      across launches and are reduced to FP64 when a request completes ---- */
typedef struct mmc_hip_arenares* mmc_hip_arena_handle;
/* An arena of nfloat field values held as `replicas` copies (1..8; the
 * kernels deposit into the copy of the XCD they run on and collect sums
 * the copies), plus a compensation buffer per copy when compensated. */
int32_t mmc_hip_arena_create(uint64_t nfloat, int32_t compensated, int32_t replicas,
                             mmc_hip_arena_handle* out);
int32_t mmc_hip_arena_release(mmc_hip_arena_handle a);
/* One persistent launch over nsim chunks; chunk i accumulates into
 * arena[fieldoff[i] .. fieldoff[i] + datalen*maxgate).  Each sim's
 * photon_offset selects which photons of its request the chunk runs. */
int32_t mmc_hip_run_chunks(mmc_hip_mesh_handle h, const mmc_hip_sim* sims, int32_t nsim,
                           mmc_hip_arena_handle arena, const uint64_t* fieldoff,
                           const mmc_hip_tune* tune, mmc_hip_result* res);
/* Launch slots (no reference counterpart): a stream with reusable device
 * buffers.  mmc_hip_run_chunks_async enqueues the same launch as
 * mmc_hip_run_chunks on the slot and returns at once; mmc_hip_slot_wait
 * waits for it and adds its results into res (the same array, kept alive
 * meanwhile); mmc_hip_slot_query returns 1 when the launch has finished, 0
 * while it runs.  With two slots a server keeps the next launch queued
 * behind the running one, so the next launch's workgroups start on the CUs
 * the running launch's last photons leave idle.  One host thread per slot. */
typedef struct mmc_hip_slotres* mmc_hip_slot_handle;
int32_t mmc_hip_slot_create(mmc_hip_slot_handle* out);
int32_t mmc_hip_slot_release(mmc_hip_slot_handle s);
int32_t mmc_hip_run_chunks_async(mmc_hip_mesh_handle h, const mmc_hip_sim* sims, int32_t nsim,
                                 mmc_hip_arena_handle arena, const uint64_t* fieldoff,
                                 const mmc_hip_tune* tune, mmc_hip_result* res, mmc_hip_slot_handle slot);
int32_t mmc_hip_slot_query(mmc_hip_slot_handle slot);
int32_t mmc_hip_slot_wait(mmc_hip_slot_handle slot, mmc_hip_result* res);
/* FP64 reduction of arena[off, off+n): adds it into host (if not NULL),
 * returns its sum, and zeroes the region for reuse. */
int32_t mmc_hip_arena_collect(mmc_hip_arena_handle a, uint64_t off, uint64_t n, double* host, double* sum);

/* ---- allocation-free launch (no counterpart in the reference) --------------
 * For callers that cannot allocate device memory: code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc.  Runs the
 * same launch as mmc_hip_run_chunks (same arguments, same results) but
 * allocates and frees nothing: the launch's own device buffers (simulation
 * tables, energy tallies, counters, work queue; detector, photon, trajectory
 * and replay buffers when the simulations ask for them) are windows of a
 * static device array of MMC_NOALLOC_SCRATCH bytes that each device keeps
 * until the process ends.  The caller passes the inputs (a resident mesh)
 * and the outputs (the arena and res).  A launch whose buffers exceed the
 * array fails with a message and runs nothing.  The array is shared by every
 * call on a device and keeps the last call's contents; calls on one device
 * are serialized by the library, and this entry point's launches must not
 * overlap other work of the caller that uses it on that device.  The first
 * call on a device looks up the array's address and copies the host C
 * library agreement tables (no allocation): run one call before capturing
 * a graph. */
#define MMC_NOALLOC_SCRATCH (8u << 20)
int32_t mmc_hip_run_chunks_noalloc(mmc_hip_mesh_handle h, const mmc_hip_sim* sims, int32_t nsim,
                                   mmc_hip_arena_handle arena, const uint64_t* fieldoff,
                                   const mmc_hip_tune* tune, mmc_hip_result* res);

#ifdef __cplusplus
}
#endif
#endif
