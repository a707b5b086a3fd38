/*
 * mmc_run_hip.c -- the HIP backend behind the reference's backend boundary.
 *
 * This file belongs in the reference's source tree (src/mmc_run_hip.c): it
 * includes the reference's own headers and gives it a fourth backend,
 *     void mmc_run_hip(mcconfig* cfg, tetmesh* mesh, raytracer* tracer)
 * with the shape of mmc_run_cl (src/mmc_cl_host.h:126) and mmc_run_mp
 * (src/mmc_host.h:54), dispatched from src/mmc.c:82-103 after mmc_prep.  It
 * translates the prepared mesh and the configuration into the plain C ABI of
 * libmmc_hip.so (include/mmc_hip.h), runs the transport on the GPU and
 * writes the results back where mmc_run_cl leaves them (src/mmc_cl_host.c:
 * 713-714, 887-902, 961-1044, 1482-1500): the field accumulated into
 * cfg->exportfield (mesh->weight by default), RF imaginary part into
 * cfg->exportadjoint, detected records into cfg->exportdetected and their
 * seeds into cfg->exportseed, cfg->detectedcount, cfg->runtime, the history
 * header, the normalization through mesh_normalize, and the standalone
 * writers mesh_saveweight / mcx_savedetphoton.
 *
 * Build (in the reference tree, with the reference's sse target flags):
 *     gcc -c -O3 -DMMC_USE_SSE -DMMC_XORSHIFT -I<this repository>/include mmc_run_hip.c
 * and link the mmc executable with -lmmc_hip.  tests/test_adapter.py
 * compiles this file against the reference's headers.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "mmc_utils.h"
#include "mmc_mesh.h"
#include "mmc_hip.h"
#include "mmc_hip_host.h"

/* the branch-less Badouel tracer keeps 4 float4 rows per element in
 * tracer->n ({nx[4]}, {ny[4]}, {nz[4]}, {d[4]}), the layout of
 * mmc_hip_mesh.planes; float3 is float4 only in the SSE / OpenCL builds
 * (src/mmc_vector_types.h:70-80) */
typedef char mmc_hip_float3_is_16_bytes[(sizeof(float3) == 16) ? 1 : -1];
typedef char mmc_hip_FLOAT3_is_12_bytes[(sizeof(FLOAT3) == 12) ? 1 : -1];
typedef char mmc_hip_medium_is_4_floats[(sizeof(medium) == 16) ? 1 : -1];
typedef char mmc_hip_float4_is_16_bytes[(sizeof(float4) == 16) ? 1 : -1];

/* values of one run's field: voxels of the dual grid, nodes or elements,
 * times gates and patterns */
static int64_t mmc_hip_field_values(const mcconfig* cfg, const tetmesh* mesh, int srcnum) {
    if (cfg->method == rtBLBadouelGrid) {
        float nmin[3];
        int32_t dim[3];

        if (mmc_host_grid((const float*)mesh->node, mesh->nn, cfg->steps.x, nmin, dim) != 0) {
            return 0;
        }

        return (int64_t)dim[0] * dim[1] * dim[2] * cfg->maxgate;
    }

    return (int64_t)(cfg->basisorder ? mesh->nn : mesh->ne) * cfg->maxgate * srcnum;
}

void mmc_run_hip(mcconfig* cfg, tetmesh* mesh, raytracer* tracer) {
    const int srcnum = cfg->srcnum > 1 ? cfg->srcnum : 1;
    const int rf = cfg->omega > 0.f;
    const int savevolume = (cfg->issave2pt || cfg->exportfield) ? 1 : 0;
    float* planes = NULL;
    double* imag = NULL;
    mmc_hip_mesh m;
    mmc_hip_sim s;
    mmc_hip_result r;

    memset(&m, 0, sizeof(m));
    memset(&s, 0, sizeof(s));
    memset(&r, 0, sizeof(r));

    /* ---- the prepared mesh (mmc_prep: tracer_build, mesh_getvolume, ...) ---- */
    m.nn = mesh->nn;
    m.ne = mesh->ne;
    m.prop = mesh->prop;
    m.nf = mesh->nf;
    m.node = (const float*)mesh->node;
    m.elem = mesh->elem;
    m.type = mesh->type;
    m.facenb = mesh->facenb;
    m.med = (const float*)mesh->med;
    m.srcelem = mesh->srcelem;
    m.srcelemlen = mesh->srcelemlen;
    m.isextdet = cfg->isextdet;

    if (cfg->method == rtBLBadouel || cfg->method == rtBLBadouelGrid) {
        m.planes = (const float*)tracer->n;
    } else {
        /* Plucker keeps edge data in the tracer; the face planes (used for
         * reflection) are formed here */
        planes = (float*)malloc(sizeof(float) * 16 * (size_t)mesh->ne);

        if (!planes) {
            MMC_ERROR(-1, "mmc_run_hip: out of memory");
        }

        mmc_host_planes(mesh->ne, (const float*)mesh->node, mesh->elem, planes);
        m.planes = planes;
    }

    /* ---- the simulation (mcconfig) ---- */
    s.nphoton = (int64_t)cfg->nphoton;
    s.seed = (uint64_t)(uint32_t)cfg->seed;
    s.photon_offset = 0;
    s.method = cfg->method == rtBLBadouel ? MMC_RT_BLBADOUEL
               : cfg->method == rtBLBadouelGrid ? MMC_RT_BLBADOUELGRID
               : cfg->method;
    s.basisorder = cfg->basisorder;
    s.outputtype = cfg->outputtype;
    s.isreflect = cfg->isreflect;
    s.isspecular = cfg->isspecular;
    s.voidtime = cfg->voidtime;
    s.srctype = cfg->srctype;
    s.srcnum = srcnum;
    s.e0 = cfg->e0;
    memcpy(s.bary0, &cfg->bary0, sizeof(float) * 4);
    memcpy(s.srcpos, &cfg->srcpos, sizeof(float) * 4);
    memcpy(s.srcdir, &cfg->srcdir, sizeof(float) * 4);
    memcpy(s.srcparam1, &cfg->srcparam1, sizeof(float) * 4);
    memcpy(s.srcparam2, &cfg->srcparam2, sizeof(float) * 4);
    s.srcpattern = cfg->srcpattern;
    s.tstart = cfg->tstart;
    s.tend = cfg->tend;
    s.tstep = cfg->tstep;
    s.maxgate = cfg->maxgate;
    s.nout = cfg->nout;
    s.minenergy = cfg->minenergy;
    s.roulettesize = cfg->roulettesize;
    s.issavedet = cfg->issavedet;
    s.issaveexit = cfg->issaveexit;
    s.detnum = cfg->detnum;
    s.detpos = (const float*)cfg->detpos;
    s.maxdetphoton = cfg->maxdetphoton;
    s.gridstep = cfg->steps.x;
    s.unitinmm = cfg->unitinmm;
    s.debuglevel = (int32_t)(cfg->debuglevel & dlTraj);
    s.omega = cfg->omega;
    s.ismomentum = cfg->ismomentum;

    if (cfg->seed == SEED_FROM_FILE && cfg->photonseed) {
        /* replay: the saved xorshift128+ states (RandType[RAND_BUF_LEN] = 16 bytes) */
        s.photonseed = (const uint64_t*)cfg->photonseed;
        s.replayweight = cfg->replayweight;
        s.replaytime = cfg->replaytime;
    }

    /* ---- outputs ---- */
    if (savevolume && cfg->exportfield == NULL) {
        cfg->exportfield = mesh->weight;
    }

    r.field = cfg->exportfield;

    if (rf && cfg->exportfield) {
        const int64_t flen = mmc_hip_field_values(cfg, mesh, srcnum);
        imag = (double*)calloc((size_t)(flen > 0 ? flen : 1), sizeof(double));
        r.field_im = imag;
    }

    if (cfg->issavedet) {
        r.reclen = (2 + (cfg->ismomentum > 0)) * mesh->prop + (cfg->issaveexit > 0) * 6 + 2;
        r.detdata = (float*)calloc((size_t)cfg->maxdetphoton * r.reclen, sizeof(float));
        r.detphoton = (int64_t*)calloc((size_t)cfg->maxdetphoton, sizeof(int64_t));
    }

    if (srcnum > 1) {
        r.srcweight = (double*)calloc(2 * (size_t)srcnum, sizeof(double));
    }

    if (cfg->debuglevel & dlTraj) {
        r.maxtraj = cfg->maxjumpdebug;
        r.traj = (float*)calloc((size_t)cfg->maxjumpdebug * MMC_TRAJ_RECLEN, sizeof(float));
    }

    if (mmc_hip_run(&m, &s, NULL, &r) != 0) {
        free(planes);
        free(imag);
        free(r.detdata);
        free(r.detphoton);
        free(r.srcweight);
        free(r.traj);
        MMC_ERROR(-1, mmc_hip_last_error());
    }

    free(planes);

    if (imag && cfg->exportadjoint) {
        const int64_t flen = mmc_hip_field_values(cfg, mesh, srcnum);
        int64_t i;

        for (i = 0; i < flen; i++) {
            cfg->exportadjoint[i] += (float)imag[i];
        }
    }

    cfg->runtime = (unsigned int)r.total_ms;

    /* ---- detected photons and their seeds ---- */
    if (cfg->issavedet) {
        const unsigned int saved = (unsigned int)(r.detected < (int64_t)cfg->maxdetphoton ? r.detected
                                                                                            : cfg->maxdetphoton);
        cfg->exportdetected = r.detdata;
        cfg->detectedcount = saved;
        cfg->his.detected = saved;

        if (cfg->issaveseed && saved > 0) {
            unsigned int i;
            cfg->exportseed = (unsigned char*)malloc((size_t)saved * 2 * sizeof(uint64_t));

            for (i = 0; cfg->exportseed && i < saved; i++) {
                mmc_hip_photon_seeds(s.seed, r.detphoton[i], 1, (uint64_t*)cfg->exportseed + 2 * (size_t)i);
            }
        }

        free(r.detphoton);
    }

    if (cfg->debuglevel & dlTraj) {
        cfg->exportdebugdata = r.traj;
        cfg->debugdatalen = (unsigned int)(r.ntraj < (int64_t)cfg->maxjumpdebug ? r.ntraj : cfg->maxjumpdebug);
    }

    /* ---- normalization (mesh_normalize per pattern, src/mmc_cl_host.c:1034-1044) ---- */
    if (cfg->isnormalized && savevolume) {
        double sum = 0.0;
        int j;

        for (j = 0; j < srcnum; j++) {
            const double etot = srcnum > 1 ? r.srcweight[j] : r.launchweight;
            const double eabs = srcnum > 1 ? r.srcweight[srcnum + j] : r.absorbweight;
            sum += mesh_normalize(mesh, cfg, (float)eabs, (float)etot, j);
        }

        cfg->his.normalizer = (float)(sum / srcnum);
    }

    free(r.srcweight);
    MMC_FPRINTF(cfg->flog, "simulated %zu photons, %.2f photon/ms, %.0f ray-tetrahedron tests\n",
                cfg->nphoton, cfg->nphoton / (r.kernel_ms > 0 ? r.kernel_ms : 1.0), r.raytet);

#ifndef MCX_CONTAINER

    if (cfg->issave2pt && cfg->parentid == mpStandalone) {
        mesh_saveweight(mesh, cfg, 0);
    }

    if (cfg->issavedet && cfg->parentid == mpStandalone && cfg->exportdetected) {
        cfg->his.totalphoton = cfg->nphoton;
        cfg->his.unitinmm = cfg->unitinmm;
        cfg->his.savedphoton = cfg->detectedcount;
        cfg->his.colcount = (2 + (cfg->ismomentum > 0)) * cfg->his.maxmedia + (cfg->issaveexit > 0) * 6 + 2;
        cfg->his.seedbyte = cfg->exportseed ? 2 * sizeof(uint64_t) : 0;
        mcx_savedetphoton(cfg->exportdetected, (void*)cfg->exportseed, cfg->detectedcount, 0, cfg);
    }

#endif
}
