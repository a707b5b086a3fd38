/*
 * mmc_oracle.h -- CPU restatement of the fangq/mmc photon-transport path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product (mmc_amd/, libmmc_hip.so) never links or calls it.
 *
 * Every function restates one reference function; the citation is given at
 * the definition in mmc_oracle.c (paths relative to the reference root).
 * The layouts below follow the reference's tetmesh/raytracer/mcconfig
 * fields (src/mmc_mesh.h:87-141, src/mmc_utils.h:210-345) reduced to plain
 * arrays so that ctypes can drive them.
 */
#ifndef MMC_ORACLE_H
#define MMC_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* method ids = enum TRTMethod (src/mmc_utils.h:97) */
#define ORC_RT_PLUCKER   0
#define ORC_RT_BLBADOUEL 3
#define ORC_RT_BLBADOUELGRID 4      /* dual-grid MMC (DMMC): voxel output */
/* source types = enum TSrcType (src/mmc_utils.h:101) */
#define ORC_ST_PENCIL      0
#define ORC_ST_ISOTROPIC   1
#define ORC_ST_CONE        2
#define ORC_ST_GAUSSIAN    3
#define ORC_ST_PLANAR      4
#define ORC_ST_PATTERN     5
#define ORC_ST_FOURIER     6
#define ORC_ST_ARCSINE     7
#define ORC_ST_DISK        8
#define ORC_ST_FOURIERX    9
#define ORC_ST_FOURIERX2D  10
#define ORC_ST_ZGAUSSIAN   11
#define ORC_ST_LINE        12
#define ORC_ST_SLIT        13
#define ORC_ST_PENCILARRAY 14
#define ORC_ST_PATTERN3D   15
#define ORC_ST_HYPERBOLOID 16
#define ORC_ST_RING        17
/* output types = enum TOutputType (src/mmc_utils.h:106) */
#define ORC_OT_FLUX    0
#define ORC_OT_FLUENCE 1
#define ORC_OT_ENERGY  2
#define ORC_OT_JACOBIAN 3     /* replay outputs, enum TOutputType (src/mmc_utils.h:106) */
#define ORC_OT_WL       4
#define ORC_OT_WP       5
/* boundary = enum TBoundary (src/mmc_utils.h:116) */
#define ORC_BC_NOREFLECT 0
#define ORC_BC_REFLECT   1
#define ORC_BC_ABSORBEXT 2
#define ORC_BC_MIRROR    3

/* termination reasons recorded per photon */
#define ORC_END_EXIT     1   /* left the mesh (eid<=0) */
#define ORC_END_LOST     2   /* no exit face after MAX_TRIAL fixes (eid=ID_UNDEFINED) */
#define ORC_END_ROULETTE 3   /* killed by Russian roulette */
#define ORC_END_STEPLIMIT 4  /* safety bound on tests per photon (the reference has none) */
#define ORC_MAX_STEPS (1u << 26)

#define ORC_ID_UNDEFINED 0x7FFFFFFF

typedef struct {
    int nn, ne, prop, nf;
    const float* node;     /* nn*3 */
    const int* elem;       /* ne*4, 1-based */
    const int* type;       /* ne, 0..prop */
    const int* facenb;     /* ne*4, exterior faces numbered -1,-2,... */
    const float* med;      /* (prop+1)*4: mua, mus, g, n; row 0 = background */
    const float* planes;   /* ne*16: nx[4] ny[4] nz[4] d[4]  (tracer->n) */
    const float* pd;       /* ne*6*3 Plucker edge directions (tracer->d) */
    const float* pm;       /* ne*6*3 Plucker edge moments    (tracer->m) */
    const float* evol;     /* ne */
    const float* nvol;     /* nn */
    const int* srcelem;    /* srcelemlen candidate launch elements */
    int srcelemlen;
    /* wide-field detector elements (mesh label -2, relabelled prop+1 with
     * the background's optical properties in med[prop+1]) */
    int isextdet;
} orc_mesh;

typedef struct {
    int64_t nphoton;
    int method, basisorder, outputtype;
    int isreflect, isspecular, voidtime, issavedet, issaveexit, ismomentum;
    int srctype, srcnum;
    int e0;
    float bary0[4];
    float srcpos[4], srcdir[4], srcparam1[4], srcparam2[4];
    const float* srcpattern;
    float tstart, tend, tstep;
    int maxgate;
    float nout, minenergy, roulettesize;
    int detnum;
    const float* detpos;              /* detnum*4: x,y,z,r */
    uint32_t seed;
    const uint64_t* photonseed;       /* nphoton*2 xorshift128+ states, or NULL */
    const float* replayweight;        /* replay J/L/P: detected weight per photon, or NULL */
    const float* replaytime;          /* replay L/P: detected time of flight per photon, or NULL */
    int nthreads;                     /* <=0: OpenMP default */
    int64_t recfirst, reccount;       /* photon-record window [recfirst, recfirst+reccount) */
    /* dual grid (ORC_RT_BLBADOUELGRID): voxel edge cfg->steps.x, grid origin
     * mesh->nmin and cfg->dim from orc_grid; unitinmm for the normalization */
    float gridstep, unitinmm;
    float gridmin[3];
    int griddim[3];
    /* RF forward (src/mmc_core.cl, IS_RF; GPU-only in the reference):
     * modulation in rad/s, > 0 for complex weights (Badouel outputs); off
     * for replays (replayweight given), as in the reference */
    float omega;
} orc_cfg;

typedef struct {
    float p[3]; float w;
    float v[3]; float t;
    int32_t eid, reason, nraytet, detid;
    float eabs, rgain, w0, nscat;
} orc_photon;

typedef struct {
    double* weight;          /* datalen*maxgate, accumulated with += */
    double launchweight, absorbweight;
    double raytet;
    float* detdata;          /* maxdet * reclen, optional */
    int64_t* detphoton;      /* maxdet photon ids (test bookkeeping), optional */
    int maxdet, detected, reclen;
    orc_photon* rec;         /* reccount records, optional */
    double runtime_ms;
    int error;
    double* weight_im;       /* RF: imaginary part of the field, or NULL */
    int32_t* ties;           /* reccount x 2 per recorded photon: tie-decided ray-tet tests, fixphoton
                                nudges; or NULL */
    double* srcweight;       /* srcnum > 1: launched energy of each pattern, then absorbed energy of
                                each pattern (2*srcnum, +=), or NULL */
} orc_out;

/* RNG (src/mmc_rand_xorshift128p.c:55-83, src/mmc_rand_common.h:50-51) */
float orc_rand_next(uint64_t t[2]);
float orc_rand_scatlen(uint64_t t[2]);
void  orc_rand_seed(const uint32_t seed[4], uint64_t t[2]);

/* mesh preparation */
void orc_getvolume(int nn, const float* node, int ne, int* elem, const int* type,
                   float* evol, float* nvol);
void orc_getfacenb(int ne, const int* elem, int* facenb);
void orc_tracer_build(int ne, const float* node, const int* elem,
                      float* planes, float* pd, float* pm);
int  orc_barycentric(const orc_mesh* m, int e0, const float srcpos[3], float bary[4]);
int  orc_initelem(const orc_mesh* m, const float srcpos[3], float bary[4]);
double orc_getreff(double n_in, double n_out);
void orc_surface_nvol(int nn, int ne, const int* elem, const int* type, const int* facenb_raw,
                      const float* med, int prop, int isreflect, float* nvol);
int  orc_number_exterior(int ne, int* facenb);
void orc_grid(int nn, const float* node, float step, float nmin[3], int dim[3]);

/* transport + normalization */
int    orc_run(const orc_cfg* cfg, const orc_mesh* mesh, orc_out* out);
double orc_normalize(const orc_cfg* cfg, const orc_mesh* mesh, double* weight,
                     float Eabsorb, float Etotal);

/* single-step probes used by unit tests */
void  orc_probe_raytet(const orc_mesh* m, int method, int eid, const float p[3], const float v[3],
                       float slen, float out[8]);
float orc_scatter(float g, float dir[3], uint64_t t[2]);
float orc_reflect(const orc_mesh* m, int method, float nout, int isreflect, int isspecular,
                  float dir[3], int oldeid, int* eid, int faceid, uint64_t t[2]);

#ifdef __cplusplus
}
#endif
#endif
