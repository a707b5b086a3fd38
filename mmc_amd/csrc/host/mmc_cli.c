/*
 * mmc_cli.c -- command-line front end of the MI355X photon tracer.
 *
 * Reads the reference's JSON input (Session / Forward / Domain / Mesh /
 * Optode, src/mmc_utils.c:1566-2392) or its mesh .dat files, honours the
 * photon-transport subset of its flags (src/mmc_utils.c:103-128,3814-4307),
 * prepares the mesh (include/mmc_hip_host.h), runs the HIP transport through the
 * C ABI (include/mmc_hip.h) and writes the reference's outputs:
 * <session>.dat (ascii, "%d\t%e" per node/element per gate, src/mmc_mesh.c:
 * 1726-1800), or .bin/.nii/.hdr+.img/.mc2/.tx3/.jnii/.bnii (src/mmc_utils.c:
 * 919-993, through mmc_host_save_field) and
 * <session>.mch (history header + detected-photon records,
 * src/mmc_utils.c:4309-4342).
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "mmc_hip.h"
#include "mmc_hip_host.h"
#include "mmc_json.h"

#define MAXDET 256

typedef struct {
    mmc_hip_sim sim;
    char session[256];
    char meshid[512];
    char rootdir[1024];
    char method;           /* 'p', 's' or 'g' */
    int isnormalized, issave2pt, outputformat; /* MMC_FMT_* (enum TOutputFormat order) */
    int zip;                        /* jnii/bnii: 1 zlib (default), 0 raw */
    long seed;
    float unitinmm;
    float gridstep;        /* dual-grid voxel edge (cfg->steps.x) */
    /* mesh given inline in the JSON */
    float* node;
    int nn;
    int* elem5;
    int ne;
    /* media, row 0 = background */
    float* med;
    int nmed;
    float detpos[4 * MAXDET];
    float detdir[4 * MAXDET];  /* Optode.Detector[].Dir: adjoint sources' directions (w = focus) */
    int ndetdir;
    float detparam1[4], detparam2[4]; /* area detector (Optode.Detector Param1/Param2): -x 2 images */
    char detpatternfile[1024];        /* .inp: detector pattern weighting the L/P replays */
    char adjoint;              /* -O a|d|u|v|w|q: adjoint Jacobian output, 0 if none */
    int adjointmode;           /* mesh mode: 0 linear-basis integrals, 1 nodal approximation */
    int nodalquarter;          /* -M S -C 1: element values split in quarters to nodes (the
                                  reference GPU backend) instead of exit-face thirds (its CPU path) */
    float* pattern;
    size_t npattern;
    int gpuid;
    /* replay (-E file.mch) and seed saving (-q 1) */
    char seedfile[1024];
    int issaveseed, replaydet;
    /* --dumpjson [file|-]: write the settings as JSON and exit */
    int dumpjson;
    char jsonfile[1024];
    /* -D M: photon trajectories, at most maxjumpdebug records */
    int savetraj;
    long long maxjumpdebug;
    uint64_t* rseeds;
    float *rweight, *rtime;
} cli_cfg;

static void die(const char* fmt, const char* arg) {
    fprintf(stderr, "MMC ERROR: ");
    fprintf(stderr, fmt, arg ? arg : "");
    fprintf(stderr, "\n");
    exit(1);
}

static void debug_flags(cli_cfg* c, const char* v);

static void defaults(cli_cfg* c) {
    memset(c, 0, sizeof(*c));
    /* mcx_initcfg, src/mmc_utils.c:207-363 */
    c->sim.nphoton = 0;
    c->seed = 0x623F9A9E;
    c->sim.isreflect = 1;
    c->isnormalized = 1;
    c->issave2pt = 1;
    c->sim.basisorder = 1;
    /* the GPU builds of the reference default to the dual grid
     * (src/mmc_utils.c:241-243) */
    c->method = 'g';
    c->gridstep = 1.f;
    c->maxjumpdebug = 10000000;       /* src/mmc_utils.c:342 */
    c->sim.minenergy = 1e-6f;
    c->sim.roulettesize = 10.f;
    c->sim.nout = 1.f;
    c->unitinmm = 1.f;
    c->sim.outputtype = MMC_OUT_FLUX;
    c->sim.voidtime = 1;
    c->sim.maxdetphoton = 1000000;
    c->sim.srcnum = 1;
    c->zip = 1;                      /* zipid = zmZlib, src/mmc_utils.c:236 */
    snprintf(c->session, sizeof(c->session), "default");
}

static char* slurp(const char* path) {
    FILE* fp = fopen(path, "rb");
    if (!fp) return NULL;
    fseek(fp, 0, SEEK_END);
    long n = ftell(fp);
    fseek(fp, 0, SEEK_SET);
    char* buf = (char*)malloc((size_t)n + 1);
    if (fread(buf, 1, (size_t)n, fp) != (size_t)n) {
        fclose(fp);
        free(buf);
        return NULL;
    }
    buf[n] = 0;
    fclose(fp);
    return buf;
}

static void vec_of(const mj_node* a, float* out, int n) {
    for (int i = 0; i < n && i < mj_size(a); i++) out[i] = (float)mj_num(mj_at(a, i), 0);
}

static int keyidx(const char* s, const char* const* table) {
    for (int i = 0; table[i][0]; i++)
        if (!strcmp(s, table[i])) return i;
    return -1;
}

/* enum TSrcType names (src/mmc_utils.c:178-180, then the four ids that
 * only the MATLAB/Python front ends name) */
/* -O / Session.OutputType (src/mmc_utils.c:161): x f e j l p, and the
 * adjoint Jacobians a d u v w q, which run fluence simulations */
static void set_outputtype(cli_cfg* c, char t) {
    t |= 0x20;
    c->adjoint = strchr("aduvwq", t) ? t : 0;
    c->sim.outputtype = (t == 'f' || c->adjoint) ? MMC_OUT_FLUENCE : t == 'e' ? MMC_OUT_ENERGY
                        : t == 'j' ? MMC_OUT_JACOBIAN : t == 'l' ? MMC_OUT_WL : t == 'p' ? MMC_OUT_WP : MMC_OUT_FLUX;
}

static const char* const kSrcTypes[] = {"pencil", "isotropic", "cone", "gaussian", "planar", "pattern", "fourier",
                                        "arcsine", "disk", "fourierx", "fourierx2d", "zgaussian", "line", "slit",
                                        "pencilarray", "pattern3d", "hyperboloid", "ring", ""};

/* mcx_loadjson subset (src/mmc_utils.c:1566-2392) */
static void load_json(cli_cfg* c, const char* path) {
    char* text = slurp(path);
    if (!text) die("can not read input file %s", path);
    char err[256];
    mj_node* root = mj_parse(text, err, sizeof(err));
    free(text);
    if (!root) die("%s", err);
    const mj_node *v, *Session = mj_get(root, "Session"), *Forward = mj_get(root, "Forward"),
                      *Domain = mj_get(root, "Domain"), *Mesh = mj_get(root, "Mesh"), *Optode = mj_get(root, "Optode");
    if (!Mesh) Mesh = mj_get(root, "Shapes");
    if (Mesh) {
        if ((v = mj_get(Mesh, "MeshID"))) snprintf(c->meshid, sizeof(c->meshid), "%s", mj_str(v, ""));
        if ((v = mj_get(Mesh, "MeshNode")) && v->type == MJ_ARRAY) {
            c->nn = mj_size(v);
            c->node = (float*)malloc(sizeof(float) * 3 * (size_t)c->nn);
            int i = 0;
            for (const mj_node* r = v->child; r; r = r->next, i++) {
                if (mj_size(r) != 3) die("%s", "Each element in MeshNode must have 3 numbers");
                vec_of(r, c->node + 3 * i, 3);
            }
        }
        if ((v = mj_get(Mesh, "MeshElem")) && v->type == MJ_ARRAY) {
            c->ne = mj_size(v);
            c->elem5 = (int*)malloc(sizeof(int) * 5 * (size_t)c->ne);
            int i = 0;
            for (const mj_node* r = v->child; r; r = r->next, i++) {
                if (mj_size(r) != 5) die("%s", "Each element in MeshElem must have 5 integers");
                for (int k = 0; k < 5; k++) c->elem5[5 * i + k] = (int)mj_num(mj_at(r, k), 0);
            }
        }
        v = mj_get(Mesh, "InitElem");
        if (!v) die("%s", "InitElem must be given");
        c->sim.e0 = (int)mj_num(v, 0);
        if ((v = mj_get(Mesh, "LengthUnit"))) c->unitinmm = (float)mj_num(v, 1);
    }
    if (Domain) {
        const mj_node* meds = mj_get(Domain, "Media");
        if (meds && meds->type == MJ_ARRAY) {
            c->nmed = mj_size(meds);
            c->med = (float*)calloc(4 * (size_t)c->nmed, sizeof(float));
            int i = 0;
            for (const mj_node* m = meds->child; m; m = m->next, i++) {
                if (m->type == MJ_OBJECT) {
                    const char* names[4] = {"mua", "mus", "g", "n"};
                    for (int k = 0; k < 4; k++) {
                        const mj_node* x = mj_get(m, names[k]);
                        if (!x) die("medium is missing field %s", names[k]);
                        c->med[4 * i + k] = (float)mj_num(x, 0);
                    }
                } else if (m->type == MJ_ARRAY && mj_size(m) >= 4) {
                    vec_of(m, c->med + 4 * i, 4);
                } else {
                    die("%s", "Domain.Media must be an array of objects or of 4-element arrays");
                }
            }
        }
        if ((v = mj_get(Domain, "LengthUnit"))) c->unitinmm = (float)mj_num(v, 1);
        /* Domain.Step: the dual grid's voxel edge (src/mmc_utils.c:1791-1801) */
        if ((v = mj_get(Domain, "Step"))) {
            if (v->type == MJ_ARRAY && mj_size(v) >= 1) c->gridstep = (float)mj_num(mj_at(v, 0), 1);
            else die("%s", "Domain::Step has incorrect element numbers");
        }
        if ((v = mj_get(Domain, "MeshID"))) snprintf(c->meshid, sizeof(c->meshid), "%s", mj_str(v, ""));
    }
    if (Optode) {
        const mj_node* src = mj_get(Optode, "Source");
        if (src) {
            if ((v = mj_get(src, "Pos"))) vec_of(v, c->sim.srcpos, 4);
            if ((v = mj_get(src, "Dir"))) vec_of(v, c->sim.srcdir, 4);
            if ((v = mj_get(src, "Type"))) {
                int t = keyidx(mj_str(v, ""), kSrcTypes);
                if (t < 0) die("unsupported source type %s", mj_str(v, ""));
                c->sim.srctype = t;
            }
            if ((v = mj_get(src, "Param1"))) vec_of(v, c->sim.srcparam1, 4);
            if ((v = mj_get(src, "Param2"))) vec_of(v, c->sim.srcparam2, 4);
            if ((v = mj_get(src, "Pattern"))) {
                const mj_node* data = v->type == MJ_OBJECT ? mj_get(v, "Data") : v;
                int n = mj_size(data);
                c->pattern = (float*)calloc((size_t)(n ? n : 1), sizeof(float));
                vec_of(data, c->pattern, n);
                c->npattern = (size_t)n;
                c->sim.srcpattern = c->pattern;
            }
        }
        const mj_node* dets = mj_get(Optode, "Detector");
        if (dets && dets->type == MJ_ARRAY) {
            int i = 0;
            for (const mj_node* d = dets->child; d && i < MAXDET; d = d->next, i++) {
                vec_of(mj_get(d, "Pos"), c->detpos + 4 * i, 3);
                c->detpos[4 * i + 3] = (float)mj_num(mj_get(d, "R"), 1.0);
                if (i == 0) {
                    const mj_node* q;
                    if ((q = mj_get(d, "Param1"))) vec_of(q, c->detparam1, 4);
                    if ((q = mj_get(d, "Param2"))) vec_of(q, c->detparam2, 4);
                }
                const mj_node* dd = mj_get(d, "Dir");
                if (dd) {
                    vec_of(dd, c->detdir + 4 * i, 4);
                    c->ndetdir = i + 1;
                }
            }
            c->sim.detnum = i;
        }
    }
    if (Session) {
        if ((v = mj_get(Session, "Photons"))) c->sim.nphoton = (int64_t)mj_num(v, 0);
        c->seed = (long)mj_num(mj_get(Session, "RNGSeed"), -1);
        if ((v = mj_get(Session, "ID"))) snprintf(c->session, sizeof(c->session), "%s", mj_str(v, "default"));
        c->sim.isreflect = (int)mj_num(mj_get(Session, "DoMismatch"), c->sim.isreflect);
        c->issave2pt = (int)mj_num(mj_get(Session, "DoSaveVolume"), c->issave2pt);
        c->isnormalized = (int)mj_num(mj_get(Session, "DoNormalize"), c->isnormalized);
        c->sim.issavedet = (int)mj_num(mj_get(Session, "DoPartialPath"), c->sim.issavedet);
        c->sim.isspecular = (int)mj_num(mj_get(Session, "DoSpecular"), c->sim.isspecular);
        c->sim.issaveexit = (int)mj_num(mj_get(Session, "DoSaveExit"), c->sim.issaveexit);
        c->sim.ismomentum = (int)mj_num(mj_get(Session, "DoDCS"), c->sim.ismomentum);
        c->issaveseed = (int)mj_num(mj_get(Session, "DoSaveSeed"), c->issaveseed);
        c->sim.basisorder = (int)mj_num(mj_get(Session, "BasisOrder"), c->sim.basisorder);
        if ((v = mj_get(Session, "OutputFormat"))) {
            c->outputformat = mmc_host_format_id(mj_str(v, "ascii"));
            if (c->outputformat < 0) die("unsupported Session.OutputFormat '%s'", mj_str(v, ""));
        }
        if ((v = mj_get(Session, "RayTracer"))) c->method = mj_str(v, "p")[0] | 0x20;
        if ((v = mj_get(Session, "DebugFlag"))) {
            char num[32];
            if (v->type == MJ_NUMBER) snprintf(num, sizeof(num), "%d", (int)v->num);
            debug_flags(c, v->type == MJ_NUMBER ? num : mj_str(v, ""));
        }
        if ((v = mj_get(Session, "OutputType"))) set_outputtype(c, mj_str(v, "x")[0]);
    }
    if (Forward) {
        c->sim.tstart = (float)mj_num(mj_get(Forward, "T0"), 0);
        c->sim.tend = (float)mj_num(mj_get(Forward, "T1"), 0);
        c->sim.tstep = (float)mj_num(mj_get(Forward, "Dt"), 0);
        c->sim.nout = (float)mj_num(mj_get(Forward, "N0"), c->sim.nout);
        /* RF modulation, rad/s (src/mmc_utils.c:2342) */
        c->sim.omega = (float)mj_num(mj_get(Forward, "Omega"), 0);
    }
    mj_free(root);
}

/* mcx_loadconfig (src/mmc_utils.c:2425-2670): the legacy text input, one
 * setting per line followed by an optional comment: photons, seed, source
 * position, direction (optional w), time gates, mesh id, initial element,
 * detector count and radius, detectors (optional own radius), then
 * optionally the source type, its two parameter sets and, for a pattern,
 * the pattern file and count. */
/* the rest of the current line (the comment after a setting); 1 if read */
static int line_rest(FILE* fp, char* buf, int len) { return fgets(buf, len, fp) != NULL; }

static void load_inp(cli_cfg* c, const char* path) {
    static const char* const names[] = {"pencil", "isotropic", "cone", "gaussian", "planar", "pattern", "fourier",
                                        "arcsine", "disk", "fourierx", "fourierx2d", "zgaussian", "line", "slit",
                                        ""};
    FILE* fp = fopen(path, "rt");
    if (!fp) die("can not read input file %s", path);
    char rest[4096], word[1024];
    long long nphoton = 0;
    int seed = 0, detnum = 0;
    float detr = 0.f, v;
    float* sp = c->sim.srcpos;
    float* sd = c->sim.srcdir;
#define INP_READ(n, ...) \
    do { \
        if (fscanf(fp, __VA_ARGS__) != (n)) die("%s", "the .inp file is incomplete"); \
    } while (0)
#define INP_REST() line_rest(fp, rest, sizeof(rest))
    INP_READ(1, "%lld", &nphoton);
    INP_REST();
    INP_READ(1, "%d", &seed);
    INP_REST();
    INP_READ(3, "%f %f %f", sp, sp + 1, sp + 2);
    INP_REST();
    INP_READ(3, "%f %f %f", sd, sd + 1, sd + 2);
    if (INP_REST() && sscanf(rest, "%f", &v) == 1) sd[3] = v;
    INP_READ(3, "%f %f %f", &c->sim.tstart, &c->sim.tend, &c->sim.tstep);
    INP_REST();
    INP_READ(1, "%1023s", word);
    snprintf(c->meshid, sizeof(c->meshid), "%.511s", word);
    INP_REST();
    INP_READ(1, "%d", &c->sim.e0);
    INP_REST();
    INP_READ(2, "%d %f", &detnum, &detr);
    INP_REST();
    if (detnum < 0 || detnum > MAXDET) die("%s", "detector count out of range");
    for (int i = 0; i < detnum; i++) {
        float* d = c->detpos + 4 * i;
        INP_READ(3, "%f %f %f", d, d + 1, d + 2);
        d[3] = (INP_REST() && sscanf(rest, "%f", &v) == 1) ? v : detr;
    }
    c->sim.detnum = detnum;
    if (fscanf(fp, "%1023s", word) == 1 && word[0]) {
        int t = keyidx(word, names);
        if (t < 0) die("%s", "the specified source type is not supported");
        c->sim.srctype = t;
        INP_REST();
        float* p1 = c->sim.srcparam1;
        float* p2 = c->sim.srcparam2;
        INP_READ(4, "%f %f %f %f", p1, p1 + 1, p1 + 2, p1 + 3);
        INP_REST();
        if (fscanf(fp, "%f %f %f %f", p2, p2 + 1, p2 + 2, p2 + 3) == 4) {
            INP_REST();
            if (t == MMC_SRC_PATTERN && p1[3] * p2[3] > 0) {
                int srcnum = 1;
                if (fscanf(fp, "%1023s %d", word, &srcnum) < 1) die("%s", "missing source pattern file");
                if (srcnum < 1) die("%s", "the number of patterns cannot be smaller than 1");
                c->sim.srcnum = srcnum;
                size_t n = (size_t)(p1[3] * p2[3]) * (size_t)srcnum;
                c->pattern = (float*)calloc(n, sizeof(float));
                FILE* pf = fopen(word, "rb");
                if (!pf) {
                    char alt[4096];
                    snprintf(alt, sizeof(alt), "%s%s", c->rootdir, word);
                    pf = fopen(alt, "rb");
                }
                if (!pf) die("%s", "source pattern file can not be opened");
                if (fread(c->pattern, sizeof(float), n, pf) != n) die("%s", "source pattern file is too short");
                c->npattern = n;
                fclose(pf);
                c->sim.srcpattern = c->pattern;
            }
        }
        if (detnum == 1 && c->detpos[3] == 0.f) {
            /* a wide-field detector plane: its two axes (pixel counts in w),
             * then optionally the pattern file used by the L/P replays
             * (src/mmc_utils.c:2612-2660) */
            float* q1 = c->detparam1;
            float* q2 = c->detparam2;
            INP_READ(4, "%f %f %f %f", q1, q1 + 1, q1 + 2, q1 + 3);
            INP_REST();
            INP_READ(4, "%f %f %f %f", q2, q2 + 1, q2 + 2, q2 + 3);
            INP_REST();
            if (q1[3] * q2[3] > 0 && fscanf(fp, "%1023s", word) == 1)
                snprintf(c->detpatternfile, sizeof(c->detpatternfile), "%s", word);
        }
    }
#undef INP_READ
#undef INP_REST
    fclose(fp);
    c->sim.nphoton = (int64_t)nphoton;
    c->seed = seed;
}

/* mcx_savejdata (src/mmc_utils.c:2710-2850): the settings as a JSON input
 * with the mesh inline, readable by -f; '-' writes to stdout */
static void dump_json(const cli_cfg* c) {
    static const char otype[] = "xfejlp";
    FILE* fp = (!c->jsonfile[0] || !strcmp(c->jsonfile, "-")) ? stdout : fopen(c->jsonfile, "wt");
    if (!fp) die("can not write %s", c->jsonfile);
    const mmc_hip_sim* s = &c->sim;
    fprintf(fp, "{\n\t\"Session\": {\n\t\t\"ID\": \"%s\",\n\t\t\"Photons\": %lld,\n\t\t\"RNGSeed\": %u,\n",
            c->session, (long long)s->nphoton, (unsigned)c->seed);
    fprintf(fp, "\t\t\"DoMismatch\": %d,\n\t\t\"DoSaveVolume\": %d,\n\t\t\"DoNormalize\": %d,\n"
                "\t\t\"DoPartialPath\": %d,\n\t\t\"DoSaveExit\": %d,\n\t\t\"DoSaveSeed\": %d,\n"
                "\t\t\"DoSpecular\": %d,\n\t\t\"DoDCS\": %d,\n\t\t\"DebugFlag\": %d,\n\t\t\"BasisOrder\": %d,\n",
            s->isreflect, c->issave2pt, c->isnormalized, s->issavedet, s->issaveexit, c->issaveseed, s->isspecular,
            s->ismomentum, c->savetraj ? MMC_DL_TRAJ : 0, s->basisorder);
    const char* fmts[] = {"ascii", "bin", "nii", "hdr", "mc2", "tx3", "jnii", "bnii"};
    fprintf(fp, "\t\t\"RayTracer\": \"%c\",\n\t\t\"OutputFormat\": \"%s\",\n\t\t\"OutputType\": \"%c\"\n\t},\n",
            c->method, fmts[c->outputformat],
            c->adjoint ? c->adjoint : otype[s->outputtype >= 0 && s->outputtype <= 5 ? s->outputtype : 0]);
    fprintf(fp, "\t\"Forward\": {\n\t\t\"T0\": %.9g,\n\t\t\"T1\": %.9g,\n\t\t\"Dt\": %.9g,\n\t\t\"N0\": %.9g,\n"
                "\t\t\"Omega\": %.9g\n\t},\n",
            (double)s->tstart, (double)s->tend, (double)s->tstep, (double)s->nout, (double)s->omega);
    fprintf(fp, "\t\"Domain\": {\n\t\t\"LengthUnit\": %.9g,\n\t\t\"Step\": [%.9g, %.9g, %.9g],\n\t\t\"Media\": [",
            (double)c->unitinmm, (double)c->gridstep, (double)c->gridstep, (double)c->gridstep);
    for (int i = 0; i < c->nmed; i++)
        fprintf(fp, "%s\n\t\t\t{\"mua\": %.9g, \"mus\": %.9g, \"g\": %.9g, \"n\": %.9g}", i ? "," : "",
                (double)c->med[4 * i], (double)c->med[4 * i + 1], (double)c->med[4 * i + 2], (double)c->med[4 * i + 3]);
    fprintf(fp, "\n\t\t]\n\t},\n\t\"Mesh\": {\n\t\t\"MeshID\": \"%s\",\n\t\t\"InitElem\": %d,\n\t\t\"MeshNode\": [",
            c->meshid, s->e0);
    for (int i = 0; i < c->nn; i++)
        fprintf(fp, "%s\n\t\t\t[%.9g, %.9g, %.9g]", i ? "," : "", (double)c->node[3 * i], (double)c->node[3 * i + 1],
                (double)c->node[3 * i + 2]);
    fprintf(fp, "\n\t\t],\n\t\t\"MeshElem\": [");
    for (int i = 0; i < c->ne; i++) {
        const int* e = c->elem5 + 5 * i;
        fprintf(fp, "%s\n\t\t\t[%d, %d, %d, %d, %d]", i ? "," : "", e[0], e[1], e[2], e[3], e[4]);
    }
    fprintf(fp, "\n\t\t]\n\t},\n\t\"Optode\": {\n\t\t\"Source\": {\n\t\t\t\"Type\": \"%s\",\n",
            kSrcTypes[s->srctype >= MMC_SRC_PENCIL && s->srctype <= MMC_SRC_RING ? s->srctype : 0]);
    fprintf(fp, "\t\t\t\"Pos\": [%.9g, %.9g, %.9g],\n\t\t\t\"Dir\": [%.9g, %.9g, %.9g, %.9g],\n",
            (double)s->srcpos[0], (double)s->srcpos[1], (double)s->srcpos[2], (double)s->srcdir[0],
            (double)s->srcdir[1], (double)s->srcdir[2], (double)s->srcdir[3]);
    fprintf(fp, "\t\t\t\"Param1\": [%.9g, %.9g, %.9g, %.9g],\n\t\t\t\"Param2\": [%.9g, %.9g, %.9g, %.9g]",
            (double)s->srcparam1[0], (double)s->srcparam1[1], (double)s->srcparam1[2], (double)s->srcparam1[3],
            (double)s->srcparam2[0], (double)s->srcparam2[1], (double)s->srcparam2[2], (double)s->srcparam2[3]);
    if ((s->srctype == MMC_SRC_PATTERN || s->srctype == MMC_SRC_PATTERN3D) && c->pattern) {
        size_t n = c->npattern;
        fprintf(fp, ",\n\t\t\t\"Pattern\": [");
        for (size_t i = 0; i < n; i++) fprintf(fp, "%s%.9g", i ? ", " : "", (double)c->pattern[i]);
        fprintf(fp, "]");
    }
    fprintf(fp, "\n\t\t},\n\t\t\"Detector\": [");
    for (int i = 0; i < s->detnum; i++) {
        fprintf(fp, "%s\n\t\t\t{\"Pos\": [%.9g, %.9g, %.9g], \"R\": %.9g", i ? "," : "", (double)c->detpos[4 * i],
                (double)c->detpos[4 * i + 1], (double)c->detpos[4 * i + 2], (double)c->detpos[4 * i + 3]);
        if (i < c->ndetdir)
            fprintf(fp, ", \"Dir\": [%.9g, %.9g, %.9g, %.9g]", (double)c->detdir[4 * i], (double)c->detdir[4 * i + 1],
                    (double)c->detdir[4 * i + 2], (double)c->detdir[4 * i + 3]);
        if (i == 0 && c->detparam1[3] * c->detparam2[3] > 0)
            fprintf(fp, ", \"Param1\": [%.9g, %.9g, %.9g, %.9g], \"Param2\": [%.9g, %.9g, %.9g, %.9g]",
                    (double)c->detparam1[0], (double)c->detparam1[1], (double)c->detparam1[2], (double)c->detparam1[3],
                    (double)c->detparam2[0], (double)c->detparam2[1], (double)c->detparam2[2], (double)c->detparam2[3]);
        fprintf(fp, "}");
    }
    fprintf(fp, "\n\t\t]\n\t}\n}\n");
    if (fp != stdout) fclose(fp);
}

/* ------------------------------------------------------------------ mesh files */
static FILE* open_mesh_file(const cli_cfg* c, const char* pattern) {
    char name[2048], path[4096];
    snprintf(name, sizeof(name), pattern, c->meshid);
    snprintf(path, sizeof(path), "%s%s", c->rootdir, name);
    FILE* fp = fopen(path, "rt");
    if (!fp) fp = fopen(name, "rt");
    return fp;
}

static void load_mesh_files(cli_cfg* c) {
    int a, b;
    FILE* fp = open_mesh_file(c, "node_%s.dat");
    if (!fp) die("can not open node file for mesh %s", c->meshid);
    if (fscanf(fp, "%d %d", &a, &c->nn) != 2 || c->nn <= 0) die("%s", "node file has wrong format");
    c->node = (float*)malloc(sizeof(float) * 3 * (size_t)c->nn);
    for (int i = 0; i < c->nn; i++)
        if (fscanf(fp, "%d %f %f %f", &a, c->node + 3 * i, c->node + 3 * i + 1, c->node + 3 * i + 2) != 4)
            die("%s", "node file has wrong format");
    fclose(fp);
    fp = open_mesh_file(c, "elem_%s.dat");
    if (!fp) die("can not open element file for mesh %s", c->meshid);
    if (fscanf(fp, "%d %d", &a, &c->ne) != 2 || c->ne <= 0) die("%s", "element file has wrong format");
    c->elem5 = (int*)malloc(sizeof(int) * 5 * (size_t)c->ne);
    for (int i = 0; i < c->ne; i++) {
        int* e = c->elem5 + 5 * i;
        if (fscanf(fp, "%d %d %d %d %d %d", &a, e, e + 1, e + 2, e + 3, e + 4) != 6)
            die("%s", "element file has wrong format");
    }
    fclose(fp);
    if (!c->med) {
        fp = open_mesh_file(c, "prop_%s.dat");
        if (!fp) die("can not open property file for mesh %s", c->meshid);
        if (fscanf(fp, "%d %d", &a, &b) != 2 || b <= 0) die("%s", "property file has wrong format");
        c->nmed = b + 1;
        c->med = (float*)calloc(4 * (size_t)c->nmed, sizeof(float));
        for (int i = 1; i <= b; i++)
            if (fscanf(fp, "%d %f %f %f %f", &a, c->med + 4 * i, c->med + 4 * i + 1, c->med + 4 * i + 2,
                       c->med + 4 * i + 3) != 5)
                die("%s", "property file has wrong format");
        fclose(fp);
    }
}

/* mesh_loadseedfile (src/mmc_mesh.c:815-898): the saved photons of detector
 * replaydet (0: all) with their seeds; for the J/L/P outputs (and any
 * replaydet) the replay weight w0 * prod exp(-mua_j * col_j * unitinmm) and
 * time sum n_j * col_j * R_C0 over record columns j = 2..maxmedia+1 and media
 * 1..maxmedia, in float as the reference computes them. */
static void load_seedfile(cli_cfg* c, int prop) {
    FILE* fp = fopen(c->seedfile, "rb");
    if (!fp) die("can not open the history file %s", c->seedfile);
    mmc_mch_header his;
    if (fread(&his, sizeof(his), 1, fp) != 1 || memcmp(his.magic, "MCXH", 4))
        die("%s is not a detected-photon history file", c->seedfile);
    if (his.savedphoton == 0 || his.seedbyte == 0) die("%s holds no photon seeds", c->seedfile);
    if (his.seedbyte != 16) die("%s", "only 16-byte xorshift128+ seeds can be replayed");
    if ((int)his.maxmedia != prop) die("%s", "the history file was generated with a different media setting");
    size_t n = his.savedphoton, col = his.colcount;
    float* ppath = (float*)malloc(sizeof(float) * n * col);
    uint64_t* seeds = (uint64_t*)malloc(sizeof(uint64_t) * 2 * n);
    if (fread(ppath, sizeof(float) * col, n, fp) != n || fread(seeds, 16, n, fp) != n)
        die("error when reading %s", c->seedfile);
    fclose(fp);
    c->rseeds = (uint64_t*)malloc(sizeof(uint64_t) * 2 * n);
    c->rweight = (float*)malloc(sizeof(float) * n);
    c->rtime = (float*)malloc(sizeof(float) * n);
    /* a patterned wide-field detector weights the L/P replays by the
     * pattern at each photon's exit pixel (src/mmc_mesh.c:871-875) */
    float* detpat = NULL;
    if (c->detpatternfile[0] && (c->sim.outputtype == MMC_OUT_WL || c->sim.outputtype == MMC_OUT_WP) &&
        c->detparam1[3] * c->detparam2[3] > 0) {
        const size_t np = (size_t)c->detparam1[3] * (size_t)c->detparam2[3];
        FILE* pf = fopen(c->detpatternfile, "rb");
        detpat = (float*)calloc(np, sizeof(float));
        if (!pf || !detpat || fread(detpat, sizeof(float), np, pf) != np)
            die("can not read the detector pattern %s", c->detpatternfile);
        fclose(pf);
    }
    const float R_C0 = 3.335640951981520e-12f; /* src/mmc_const.h:45 */
    size_t m = 0;
    for (size_t i = 0; i < n; i++) {
        const float* r = ppath + i * col;
        if (c->replaydet != 0 && c->replaydet != (int)r[0]) continue;
        c->rseeds[2 * m] = seeds[2 * i];
        c->rseeds[2 * m + 1] = seeds[2 * i + 1];
        float w = r[col - 1], t = 0.f;
        if (detpat && mmc_host_detweight(&w, r, 1, (int32_t)col, c->detpos, c->detparam1, c->detparam2, detpat))
            die("%s", "photon location not within the detection plane");
        for (int j = 2; j < prop + 2; j++) w *= expf(-c->med[4 * (j - 1)] * r[j] * his.unitinmm);
        for (int j = 2; j < prop + 2; j++) t += c->med[4 * (j - 1) + 3] * r[j] * R_C0;
        c->rweight[m] = w;
        c->rtime[m] = t;
        m++;
    }
    free(detpat);
    free(ppath);
    free(seeds);
    if (m == 0) die("%s", "no saved photon of the replayed detector");
    c->sim.nphoton = (int64_t)m;
    c->sim.photon_offset = 0;
    c->sim.photonseed = c->rseeds;
    c->sim.replayweight = c->rweight;
    c->sim.replaytime = c->rtime;
    if (c->sim.outputtype >= MMC_OUT_JACOBIAN || c->replaydet > 0) c->sim.minenergy = 0.f;
}

static void usage(void) {
    printf(
        "mmc (MI355X HIP backend) -- mesh Monte Carlo photon transport\n"
        "usage: mmc -f input.json [options]\n"
        "  -f/--input FILE       JSON input (Session/Forward/Domain/Mesh/Optode)\n"
        "  -n/--photon N         photons to simulate\n"
        "  -E/--seed N|file.mch  RNG seed (negative: time-based), or replay the saved photons of an .mch\n"
        "  -q/--saveseed [0|1]   save the seeds of detected photons in the .mch (for replay)\n"
        "  -D/--debug FLAGS      M: save photon trajectories to <session>.mct (other letters of the\n"
        "                        reference only change its logging and are accepted)\n"
        "  --maxjumpdebug N      trajectory positions kept (default 10000000)\n"
        "  --dumpjson [FILE|-]   write the settings (mesh inline) as JSON input and exit\n"
        "  -P/--replaydet N      replay only the photons of detector N (0: all)\n"
        "  -s/--session NAME     output file prefix\n"
        "  -M/--method [G|P|S]   dual grid (voxel output, default), Plucker (nodal, -C 1) or\n"
        "                        branch-less Badouel (element with -C 0, nodal with -C 1)\n"
        "  --gridsize D          dual-grid voxel edge in mesh units (default 1)\n"
        "  -C/--basisorder [0|1] element (0) or nodal (1) fluence\n"
        "  -O/--outputtype [X|F|E|J|L|P|A|D|U|V|W|Q] flux, fluence, energy deposit; replay:\n"
        "                        Jacobian, weighted path length, weighted scattering count;\n"
        "                        adjoint Jacobians (detectors need Dir): mua, D, mus, musp,\n"
        "                        mua and D, mua and musp -> <session>_jmua, _jd, _jmus, _jmusp\n"
        "  --adjointmode [0|1]   mesh-mode J_mua: linear-basis integrals (0) or nodal approx. (1)\n"
        "  --nodalsplit [face|quarter]  -M S -C 1: a third of each step to the exit face's nodes\n"
        "                        (the reference CPU path, default) or element values in quarters\n"
        "                        to their nodes (the reference GPU backend)\n"
        "  -F/--outputformat [ascii|bin|nii|hdr|mc2|tx3|jnii|bnii]\n"
        "  -Z/--zip [zlib|base64]   jnii/bnii payload: zlib-compressed (default) or raw\n"
        "  -b/--reflect [0|1|2|3] boundary: no reflect, Fresnel, absorb exterior, mirror\n"
        "  -U/--normalize [0|1]  normalize the output\n"
        "  -S/--save2pt [0|1]    save the fluence field\n"
        "  -d/--savedet [0|1]    save detected photons (.mch)\n"
        "  -x/--saveexit [0|1]   save exit position and direction of detected photons\n"
        "  -m/--momentum [0|1]   save the momentum transfer per medium of detected photons\n"
        "  -H/--maxdetphoton N   detected-photon buffer size\n"
        "  -e/--minenergy W      roulette threshold\n"
        "  -V/--specular [0|1]   specular reflection at launch\n"
        "  -k/--voidtime [0|1]   count time in void elements\n"
        "  -u/--unitinmm U       mesh length unit in mm\n"
        "  -G/--gpu 1            the GPU (one GPU only)\n"
        "  -c/--compute hip      backend (hip only)\n");
}

/* -D / Session.DebugFlag: letters of debugflag[] (src/mmc_utils.c:139) or a
 * number; M (dlTraj) records trajectories, the rest select log output */
static void debug_flags(cli_cfg* c, const char* v) {
    if (v[0] >= '0' && v[0] <= '9') {
        if (atoi(v) & MMC_DL_TRAJ) c->savetraj = 1;
        return;
    }
    for (const char* p = v; *p; p++) {
        char u = (char)(*p & ~0x20);
        if (!strchr("SCBWDIOXATRPEM", u)) die("unknown debug flag '%s'", v);
        if (u == 'M') c->savetraj = 1;
    }
}

static const char* opt_value(int argc, char** argv, int* i) {
    if (*i + 1 >= argc) die("missing value for %s", argv[*i]);
    return argv[++*i];
}

/* Adjoint Jacobians (-O a|d|u|v|w|q; src/mmc_utils.c:3755-3800,
 * src/mmc_cl_host.c:1117-1462, src/mmc_mesh.c:1802-1960): the source, then
 * each detector launched from its position along Optode.Detector[].Dir (a
 * disk of its radius), each run as its own fluence simulation with an equal
 * share of the photons and normalized alone; the Jacobians of the normalized
 * fields are written as <session>_jmua / _jd (_jmus, _jmusp: the same J_D
 * product under the reference's names), values [s][d][cell] with the
 * imaginary block after the real one for RF. */
static int run_adjoint(cli_cfg* c, const mmc_hip_mesh* mesh, const int32_t* elem, const int32_t* type,
                       const float* evol, const float* nvol, int isgrid, const int32_t griddim[3], int datalen) {
    const int nd = c->sim.detnum, ns = 1, nslot = ns + nd;
    if (nd < 1 || c->ndetdir < nd) die("%s", "adjoint outputs need detectors with directions (Optode.Detector[].Dir)");
    if (!isgrid && c->sim.basisorder != 1)
        die("%s", "mesh-mode adjoint Jacobians need nodal fluence (-M P -C 1) or use the dual grid (-M G)");
    if (c->sim.srctype == MMC_SRC_PATTERN && c->sim.srcnum > 1) die("%s", "adjoint outputs take one source pattern");
    const int isrf = c->sim.omega > 0.f;
    const size_t cells = (size_t)datalen, per = cells * (size_t)c->sim.maxgate;
    float* f = (float*)calloc(per * nslot, sizeof(float));
    float* fi = isrf ? (float*)calloc(per * nslot, sizeof(float)) : NULL;
    double* field = (double*)calloc(per, sizeof(double));
    double* field_im = isrf ? (double*)calloc(per, sizeof(double)) : NULL;
    if (!f || !field || (isrf && (!fi || !field_im))) die("%s", "can not allocate the adjoint fields");
    const int64_t nph = c->sim.nphoton / nslot > 0 ? c->sim.nphoton / nslot : 1;
    for (int k = 0; k < nslot; k++) {
        mmc_hip_sim s = c->sim;
        s.nphoton = nph;
        s.issavedet = 0;
        s.issaveexit = 0;
        s.detnum = 0;
        s.detpos = NULL;
        if (k >= ns) {
            const float* dp = c->detpos + 4 * (k - ns);
            const float* dd = c->detdir + 4 * (k - ns);
            memcpy(s.srcpos, dp, 3 * sizeof(float));
            s.srcpos[3] = 1.f;
            memcpy(s.srcdir, dd, 4 * sizeof(float));
            memset(s.srcparam1, 0, sizeof(s.srcparam1));
            memset(s.srcparam2, 0, sizeof(s.srcparam2));
            s.srcparam1[0] = dp[3];
            s.srctype = dp[3] > 0.f ? MMC_SRC_DISK : MMC_SRC_PENCIL;
            s.srcpattern = NULL;
            s.srcnum = 1;
            float b[4] = {0.f, 0.f, 0.f, 0.f};
            s.e0 = mmc_host_initelem(c->ne, c->node, elem, s.srcpos, b);
            if (s.e0 <= 0) die("%s", "a detector lies outside the mesh (adjoint source)");
            memcpy(s.bary0, b, sizeof(b));
        }
        mmc_hip_result res;
        memset(&res, 0, sizeof(res));
        memset(field, 0, per * sizeof(double));
        if (field_im) memset(field_im, 0, per * sizeof(double));
        res.field = field;
        res.field_im = field_im;
        if (mmc_hip_run(mesh, &s, NULL, &res) != 0) die("%s", mmc_hip_last_error());
        const double k_norm = isgrid ? mmc_host_normalize_grid(&s, datalen, field, res.absorbweight,
                                                               res.launchweight, 0)
                                     : mmc_host_normalize(&s, c->nn, c->ne, elem, type, c->med, evol, nvol, field,
                                                          res.absorbweight, res.launchweight);
        for (size_t i = 0; i < per; i++) f[k * per + i] = (float)field[i];
        if (fi)
            for (size_t i = 0; i < per; i++) fi[k * per + i] = (float)(field_im[i] * k_norm);
        printf("adjoint slot %d of %d: %lld photons, %.2f photon/ms\n", k + 1, nslot, (long long)nph,
               (double)nph / res.total_ms);
    }
    const char t = c->adjoint;
    const int want_mua = (t == 'a' || t == 'w' || t == 'q'), want_d = (t != 'a');
    const size_t nout = cells * (size_t)(ns * nd) * (isrf ? 2 : 1);
    float* jm = want_mua ? (float*)calloc(nout, sizeof(float)) : NULL;
    float* jd = want_d ? (float*)calloc(nout, sizeof(float)) : NULL;
    int rc;
    if (isgrid) {
        rc = mmc_hip_adjoint_grid(f, fi, griddim, c->sim.maxgate, ns, nd, c->unitinmm, jm, jd);
    } else {
        float* ddd = NULL;
        if (want_d) {
            double* d64 = (double*)malloc(sizeof(double) * 10 * (size_t)c->ne);
            ddd = (float*)malloc(sizeof(float) * 10 * (size_t)c->ne);
            if (!d64 || !ddd || mmc_host_deldotdel(c->node, c->nn, elem, c->ne, evol, d64))
                die("%s", "can not form the element gradient products");
            for (size_t i = 0; i < 10 * (size_t)c->ne; i++) ddd[i] = (float)d64[i];
            free(d64);
        }
        rc = mmc_hip_adjoint_mesh(f, fi, c->nn, c->ne, elem, evol, ddd, nvol, c->sim.maxgate, ns, nd,
                                  c->adjointmode == 1 ? 1 : 0, jm, jd);
        free(ddd);
    }
    if (rc) die("adjoint Jacobians failed: %s", mmc_hip_last_error());
    const char* name_d = (t == 'u') ? "jmus" : (t == 'v' || t == 'q') ? "jmusp" : "jd";
    for (int w = 0; w < 2; w++) {
        const float* j = w == 0 ? jm : jd;
        if (!j) continue;
        char sess[400];
        snprintf(sess, sizeof(sess), "%s_%s", c->session, w == 0 ? "jmua" : name_d);
        double* out = (double*)malloc(sizeof(double) * nout);
        if (!out) die("%s", "can not allocate the Jacobian output");
        for (size_t i = 0; i < nout; i++) out[i] = j[i];
        const int nblk = ns * nd * (isrf ? 2 : 1);
        rc = isgrid ? mmc_host_save_grid(sess, c->outputformat, out, griddim, c->gridstep, nblk, 1, c->sim.tstep,
                                         MMC_OUT_FLUENCE, c->zip)
                    : mmc_host_save_field(sess, c->outputformat, out, datalen, nblk, 1, c->sim.tstep,
                                          MMC_OUT_FLUENCE, c->zip);
        free(out);
        if (rc) die("can not write %s", sess);
        printf("saved %s (%d source-detector pairs%s)\n", sess, ns * nd, isrf ? ", real then imaginary" : "");
    }
    free(jm);
    free(jd);
    free(f);
    free(fi);
    free(field);
    free(field_im);
    return 0;
}

int main(int argc, char** argv) {
    cli_cfg c;
    defaults(&c);
    const char* input = NULL;
    for (int i = 1; i < argc; i++)
        if (!strcmp(argv[i], "-f") || !strcmp(argv[i], "--input")) input = (i + 1 < argc) ? argv[i + 1] : NULL;
    if (argc == 1 || !input) {
        for (int i = 1; i < argc; i++)
            if (!strcmp(argv[i], "-h") || !strcmp(argv[i], "--help")) {
                usage();
                return 0;
            }
        usage();
        return input ? 0 : 1;
    }
    {
        const char* slash = strrchr(input, '/');
        if (slash) snprintf(c.rootdir, sizeof(c.rootdir), "%.*s/", (int)(slash - input), input);
    }
    /* JSON input, else the legacy .inp text (mcx_readconfig, src/mmc_utils.c:1475-1500) */
    if (strstr(input, ".json") || strstr(input, ".JSON")) load_json(&c, input);
    else load_inp(&c, input);
    /* flags given on the command line win over the file */
    for (int i = 1; i < argc; i++) {
        const char* a = argv[i];
        if (!strcmp(a, "-f") || !strcmp(a, "--input")) i++;
        else if (!strcmp(a, "-n") || !strcmp(a, "--photon")) c.sim.nphoton = (int64_t)atof(opt_value(argc, argv, &i));
        else if (!strcmp(a, "-E") || !strcmp(a, "--seed")) {
            const char* v = opt_value(argc, argv, &i);
            size_t n = strlen(v);
            if (n > 4 && !strcmp(v + n - 4, ".mch")) snprintf(c.seedfile, sizeof(c.seedfile), "%s", v);
            else c.seed = atol(v);
        } else if (!strcmp(a, "-q") || !strcmp(a, "--saveseed")) c.issaveseed = atoi(opt_value(argc, argv, &i));
        else if (!strcmp(a, "-D") || !strcmp(a, "--debug")) debug_flags(&c, opt_value(argc, argv, &i));
        else if (!strcmp(a, "--maxjumpdebug")) c.maxjumpdebug = atoll(opt_value(argc, argv, &i));
        else if (!strcmp(a, "--dumpjson")) {
            /* --dumpjson [file]: stdout unless a file name follows */
            c.dumpjson = 1;
            if (i + 1 < argc && (argv[i + 1][0] != '-' || !strcmp(argv[i + 1], "-")))
                snprintf(c.jsonfile, sizeof(c.jsonfile), "%s", argv[++i]);
        }
        else if (!strcmp(a, "-P") || !strcmp(a, "--replaydet")) c.replaydet = atoi(opt_value(argc, argv, &i));
        else if (!strcmp(a, "-s") || !strcmp(a, "--session")) snprintf(c.session, sizeof(c.session), "%s", opt_value(argc, argv, &i));
        else if (!strcmp(a, "-M") || !strcmp(a, "--method")) c.method = opt_value(argc, argv, &i)[0] | 0x20;
        else if (!strcmp(a, "-C") || !strcmp(a, "--basisorder")) c.sim.basisorder = atoi(opt_value(argc, argv, &i));
        else if (!strcmp(a, "--gridsize")) c.gridstep = (float)atof(opt_value(argc, argv, &i));
        else if (!strcmp(a, "--adjointmode")) c.adjointmode = atoi(opt_value(argc, argv, &i));
        else if (!strcmp(a, "--nodalsplit")) {
            const char* v = opt_value(argc, argv, &i);
            if (!strcmp(v, "quarter")) c.nodalquarter = 1;
            else if (!strcmp(v, "face")) c.nodalquarter = 0;
            else die("%s", "--nodalsplit takes face or quarter");
        }
        else if (!strcmp(a, "-m") || !strcmp(a, "--momentum")) c.sim.ismomentum = atoi(opt_value(argc, argv, &i));
        else if (!strcmp(a, "-O") || !strcmp(a, "--outputtype")) {
            set_outputtype(&c, opt_value(argc, argv, &i)[0]);
        } else if (!strcmp(a, "-F") || !strcmp(a, "--outputformat")) {
            const char* f = opt_value(argc, argv, &i);
            if ((c.outputformat = mmc_host_format_id(f)) < 0) die("unsupported output format '%s'", f);
        } else if (!strcmp(a, "-Z") || !strcmp(a, "--zip")) {
            const char* z = opt_value(argc, argv, &i);
            if (!strcmp(z, "zlib")) c.zip = 1;
            else if (!strcmp(z, "base64")) c.zip = 0;
            else die("unsupported compression '%s' (zlib or base64)", z);
        } else if (!strcmp(a, "-b") || !strcmp(a, "--reflect")) c.sim.isreflect = atoi(opt_value(argc, argv, &i));
        else if (!strcmp(a, "-U") || !strcmp(a, "--normalize")) c.isnormalized = atoi(opt_value(argc, argv, &i));
        else if (!strcmp(a, "-S") || !strcmp(a, "--save2pt")) c.issave2pt = atoi(opt_value(argc, argv, &i));
        else if (!strcmp(a, "-d") || !strcmp(a, "--savedet")) c.sim.issavedet = atoi(opt_value(argc, argv, &i));
        else if (!strcmp(a, "-x") || !strcmp(a, "--saveexit")) c.sim.issaveexit = atoi(opt_value(argc, argv, &i));
        else if (!strcmp(a, "-H") || !strcmp(a, "--maxdetphoton")) c.sim.maxdetphoton = atol(opt_value(argc, argv, &i));
        else if (!strcmp(a, "-e") || !strcmp(a, "--minenergy")) c.sim.minenergy = (float)atof(opt_value(argc, argv, &i));
        else if (!strcmp(a, "-V") || !strcmp(a, "--specular")) c.sim.isspecular = atoi(opt_value(argc, argv, &i));
        else if (!strcmp(a, "-k") || !strcmp(a, "--voidtime")) c.sim.voidtime = atoi(opt_value(argc, argv, &i));
        else if (!strcmp(a, "-u") || !strcmp(a, "--unitinmm")) c.unitinmm = (float)atof(opt_value(argc, argv, &i));
        else if (!strcmp(a, "-G") || !strcmp(a, "--gpu")) c.gpuid = atoi(opt_value(argc, argv, &i));
        else if (!strcmp(a, "-c") || !strcmp(a, "--compute")) {
            const char* b = opt_value(argc, argv, &i);
            if (strcmp(b, "hip")) die("this build has only the hip backend (asked for %s)", b);
        } else if (!strcmp(a, "-h") || !strcmp(a, "--help")) {
            usage();
            return 0;
        } else die("unknown or unsupported option %s", a);
    }
    if (c.gpuid < 0) die("%s", "the CPU path (-G -1) is not part of this build");
    if (c.sim.nphoton <= 0 && !c.seedfile[0]) die("%s", "cfg.nphoton must be a positive number");
    if (c.sim.tstart > c.sim.tend || c.sim.tstep == 0.f) die("%s", "incorrect time gate settings or missing tstart/tend/tstep fields");
    /* mcx_validatecfg, src/mmc_utils.c:3500-3530 */
    if (c.sim.tstep > c.sim.tend - c.sim.tstart) c.sim.tstep = c.sim.tend - c.sim.tstart;
    c.sim.maxgate = (int)((c.sim.tend - c.sim.tstart) / c.sim.tstep + 0.5);
    c.sim.tend = c.sim.tstart + c.sim.tstep * c.sim.maxgate;
    if (c.seed < 0) c.seed = (long)time(NULL);
    c.sim.seed = (uint64_t)(uint32_t)c.seed;
    c.sim.method = (c.method == 'p')   ? MMC_RT_PLUCKER
                   : (c.method == 's') ? MMC_RT_BLBADOUEL
                   : (c.method == 'g') ? MMC_RT_BLBADOUELGRID : -1;
    if (c.sim.method < 0) die("%s", "only -M G, -M P and -M S are available in this build");
    /* the dual grid deposits per voxel (src/mmc_utils.c:3554-3556) */
    if (c.sim.method == MMC_RT_BLBADOUELGRID) {
        c.sim.basisorder = 0;
        if (!(c.gridstep > 0.f)) die("%s", "the dual-grid voxel edge (--gridsize, Domain.Step) must be positive");
    }
    c.sim.gridstep = c.gridstep;
    c.sim.unitinmm = c.unitinmm;
    c.sim.detpos = c.detpos;

    if (!c.node) load_mesh_files(&c);
    if (!c.med) die("%s", "no media given (Domain.Media or prop_<id>.dat)");
    if (c.dumpjson) {
        dump_json(&c);
        return 0;
    }
    c.med[0] = 0.f;
    c.med[1] = 0.f;
    c.med[2] = 1.f;
    c.med[3] = c.sim.nout;
    if (c.unitinmm != 1.f)
        for (int i = 1; i < c.nmed; i++) {
            c.med[4 * i] *= c.unitinmm;
            c.med[4 * i + 1] *= c.unitinmm;
        }
    int prop = c.nmed - 1;
    if (c.seedfile[0]) load_seedfile(&c, prop);
    else if (c.sim.outputtype >= MMC_OUT_JACOBIAN) die("%s", "the J/L/P outputs replay photons: give -E file.mch");

    /* mmc_prep: volumes, neighbours, surface correction, exterior faces, planes */
    int32_t* elem = (int32_t*)malloc(sizeof(int32_t) * 4 * (size_t)c.ne);
    int32_t* type = (int32_t*)malloc(sizeof(int32_t) * (size_t)c.ne);
    for (int i = 0; i < c.ne; i++) {
        for (int k = 0; k < 4; k++) elem[4 * i + k] = c.elem5[5 * i + k];
        type[i] = c.elem5[5 * i + 4];
    }
    /* wide-field detector elements (label -2) become label prop+1 and
     * replace the point detectors (mesh_srcdetelem, src/mmc_mesh.c:390-427) */
    const int isextdet = mmc_host_detelem(c.ne, type, prop) > 0;
    if (isextdet) {
        /* the detector label's medium is a copy of the background (src/mmc_mesh.c:518-520) */
        c.sim.detnum = 0;
        c.med = (float*)realloc(c.med, sizeof(float) * 4 * (size_t)(c.nmed + 1));
        memcpy(c.med + 4 * c.nmed, c.med, 4 * sizeof(float));
    }
    /* source elements (label -1): the launch search's candidates, relabelled
     * 0; the first is the default initial element (src/mmc_mesh.c:390-427) */
    int32_t* srcelem = (int32_t*)malloc(sizeof(int32_t) * (size_t)(c.ne > 0 ? c.ne : 1));
    const int nsrcelem = mmc_host_srcelem(c.ne, type, srcelem);
    if (c.sim.e0 <= 0 && nsrcelem > 0) c.sim.e0 = srcelem[0];
    /* pattern sizes the kernel will read */
    if (c.sim.srctype == MMC_SRC_PATTERN || c.sim.srctype == MMC_SRC_PATTERN3D) {
        const float* p1 = c.sim.srcparam1;
        const double need = c.sim.srctype == MMC_SRC_PATTERN
                                ? (double)(int)p1[3] * (int)c.sim.srcparam2[3] * (c.sim.srcnum > 0 ? c.sim.srcnum : 1)
                                : (double)(int)p1[0] * (int)p1[1] * (int)p1[2];
        if (!c.pattern || need < 1 || need > (double)c.npattern) {
            char msg[96];
            snprintf(msg, sizeof(msg), "%.0f values, %zu given", need, c.npattern);
            die("the source pattern needs %s", msg);
        }
    }
    if (c.sim.issavedet && c.sim.detnum == 0 && !isextdet) c.sim.issavedet = 0;
    if (!c.sim.issavedet) {
        c.sim.issaveexit = 0;
        c.sim.ismomentum = 0;
    }
    for (int i = 0; i < c.ne; i++)
        if (type[i] < 0 || type[i] > prop + isextdet) die("%s", "element labels outside the media list");
    float* evol = (float*)malloc(sizeof(float) * (size_t)c.ne);
    float* nvol = (float*)malloc(sizeof(float) * (size_t)c.nn);
    int32_t* facenb = (int32_t*)malloc(sizeof(int32_t) * 4 * (size_t)c.ne);
    float* planes = (float*)malloc(sizeof(float) * 16 * (size_t)c.ne);
    mmc_host_volume(c.nn, c.node, c.ne, elem, type, evol, nvol);
    if (mmc_host_facenb(c.nn, c.ne, elem, facenb)) die("%s", "mesh has element node ids outside 1..nn");
    if (c.sim.basisorder && c.isnormalized == 1)
        mmc_host_surface_nvol(c.nn, c.ne, elem, type, facenb, c.med, prop, c.sim.isreflect, nvol);
    int nf = mmc_host_number_exterior(c.ne, facenb);
    mmc_host_planes(c.ne, c.node, elem, planes);
    const int st = c.sim.srctype;
    if (st == MMC_SRC_PENCIL || st == MMC_SRC_ISOTROPIC || st == MMC_SRC_CONE || st == MMC_SRC_ARCSINE) {
        /* launched at srcpos: the enclosing element and barycentric weights */
        if (c.sim.e0 <= 0 || c.sim.e0 > c.ne ||
            mmc_host_barycentric(c.node, elem, c.sim.e0, c.sim.srcpos, c.sim.bary0)) {
            c.sim.e0 = mmc_host_initelem(c.ne, c.node, elem, c.sim.srcpos, c.sim.bary0);
            if (c.sim.e0 <= 0) die("%s", "initial element does not enclose the source!");
        }
    } else if (c.sim.e0 <= 0 || c.sim.e0 > c.ne) {
        /* area sources search e0 first, then the source elements */
        float b[4];
        c.sim.e0 = mmc_host_initelem(c.ne, c.node, elem, c.sim.srcpos, b);
        if (c.sim.e0 <= 0) c.sim.e0 = 1;
    }
    mmc_hip_mesh mesh = {c.nn, c.ne, prop, nf, c.node, elem, type, facenb, planes, c.med,
                         nsrcelem ? srcelem : NULL, nsrcelem, isextdet};
    int datalen = c.sim.basisorder ? c.nn : c.ne;
    float gridmin[3] = {0.f, 0.f, 0.f};
    int32_t griddim[3] = {0, 0, 0};
    const int isgrid = c.sim.method == MMC_RT_BLBADOUELGRID;
    if (isgrid) {
        if (mmc_host_grid(c.node, c.nn, c.gridstep, gridmin, griddim)) die("%s", "bad dual-grid voxel edge");
        int64_t nc = (int64_t)griddim[0] * griddim[1] * griddim[2];
        if (nc >= (1ll << 31)) die("%s", "the dual grid has more than 2^31 voxels (raise --gridsize)");
        datalen = (int)nc;
    }
    if (c.adjoint) {
        int rc = run_adjoint(&c, &mesh, elem, type, evol, nvol, isgrid, griddim, datalen);
        free(elem);
        free(type);
        free(srcelem);
        free(evol);
        free(nvol);
        free(facenb);
        free(planes);
        return rc;
    }
    double* field = (double*)calloc((size_t)datalen * c.sim.maxgate, sizeof(double));
    /* RF forward: the imaginary part of the complex fluence */
    const int isrf = c.sim.omega > 0.f && !c.rseeds;
    double* field_im = isrf ? (double*)calloc((size_t)datalen * c.sim.maxgate, sizeof(double)) : NULL;
    mmc_hip_result res;
    memset(&res, 0, sizeof(res));
    res.field = field;
    res.field_im = field_im;
    if (c.savetraj) {
        if (c.maxjumpdebug < 1) die("%s", "--maxjumpdebug must be positive");
        c.sim.debuglevel |= MMC_DL_TRAJ;
        res.maxtraj = c.maxjumpdebug;
        res.traj = (float*)calloc((size_t)c.maxjumpdebug * MMC_TRAJ_RECLEN, sizeof(float));
        if (!res.traj) die("%s", "can not allocate the trajectory buffer");
    }
    int reclen = (2 + (c.sim.ismomentum > 0)) * prop + 6 * (c.sim.issaveexit > 0) + 2;
    if (c.sim.issavedet) {
        res.detdata = (float*)calloc((size_t)c.sim.maxdetphoton * reclen, sizeof(float));
        res.detphoton = (int64_t*)calloc((size_t)c.sim.maxdetphoton, sizeof(int64_t));
        res.reclen = reclen;
    }
    char dev[256] = "";
    mmc_hip_device_info(dev, sizeof(dev), NULL);
    printf("simulating %lld photons on %s (seed %u, method %c, %d gates)\n", (long long)c.sim.nphoton, dev,
           (unsigned)c.sim.seed, c.method == 'p' ? 'P' : c.method == 's' ? 'S' : 'G', c.sim.maxgate);
    if (isgrid)
        printf("dual grid: %d x %d x %d voxels of %g from (%g, %g, %g)\n", griddim[0], griddim[1], griddim[2],
               (double)c.gridstep, (double)gridmin[0], (double)gridmin[1], (double)gridmin[2]);
    /* the reference GPU backend's nodal output under the Badouel tracer: an
     * element-output run, then quarters of each element value to its nodes */
    const int quarter = c.nodalquarter && c.sim.method == MMC_RT_BLBADOUEL && c.sim.basisorder == 1;
    double* felem = NULL;
    if (quarter) {
        if (isrf) die("%s", "RF fluence is element or grid output (-C 0 or -M G)");
        felem = (double*)calloc((size_t)c.ne * c.sim.maxgate, sizeof(double));
        if (!felem) die("%s", "can not allocate the element field");
        c.sim.basisorder = 0;
        res.field = felem;
    }
    if (mmc_hip_run(&mesh, &c.sim, NULL, &res) != 0) die("%s", mmc_hip_last_error());
    if (quarter) {
        c.sim.basisorder = 1;
        if (mmc_host_elem_to_nodes(c.ne, c.nn, elem, c.sim.maxgate, 1, felem, field))
            die("%s", "element-to-node split failed");
        free(felem);
        res.field = field;
    }
    printf("speed ...\t%.2f photon/ms, %.0f ray-tetrahedron tests (%.2f test/ms)\n",
           (double)c.sim.nphoton / res.total_ms, res.raytet, res.raytet / res.kernel_ms);
    double normalizer = 1.0;
    if (c.isnormalized) {
        mmc_hip_sim s = c.sim;
        normalizer = isgrid ? mmc_host_normalize_grid(&s, datalen, field, res.absorbweight, res.launchweight, 0)
                            : mmc_host_normalize(&s, c.nn, c.ne, elem, type, c.med, evol, nvol, field,
                                                 res.absorbweight, res.launchweight);
        /* the imaginary part takes the same factor (src/mmc_mesh.c:2262-2274) */
        if (field_im)
            for (size_t k = 0; k < (size_t)datalen * c.sim.maxgate; k++) field_im[k] *= normalizer;
    }
    printf("total simulated energy: %.0f\tabsorbed: %5.5f%%\tnormalizor=%g\n", res.launchweight,
           100.0 * res.absorbweight / res.launchweight, normalizer);
    char fname[1024];
    if (c.issave2pt) {
        int rc = isgrid ? mmc_host_save_grid(c.session, c.outputformat, field, griddim, c.gridstep, c.sim.maxgate,
                                             1, c.sim.tstep, c.sim.outputtype, c.zip)
                        : mmc_host_save_field(c.session, c.outputformat, field, datalen, c.sim.maxgate, 1,
                                              c.sim.tstep, c.sim.outputtype, c.zip);
        if (!rc && field_im) {
            /* the reference returns it to its MATLAB interface only; written
             * here as <session>_imag.<ext> */
            char sess_im[300];
            snprintf(sess_im, sizeof(sess_im), "%s_imag", c.session);
            rc = isgrid ? mmc_host_save_grid(sess_im, c.outputformat, field_im, griddim, c.gridstep, c.sim.maxgate, 1,
                                             c.sim.tstep, c.sim.outputtype, c.zip)
                        : mmc_host_save_field(sess_im, c.outputformat, field_im, datalen, c.sim.maxgate, 1,
                                              c.sim.tstep, c.sim.outputtype, c.zip);
        }
        if (rc)
            die("can not write the output of session %s", c.session);
    }
    if (c.sim.issavedet) {
        int64_t saved = res.detected < c.sim.maxdetphoton ? res.detected : c.sim.maxdetphoton;
        mmc_mch_header his;
        memset(&his, 0, sizeof(his));
        memcpy(his.magic, "MCXH", 4);
        his.version = 1;
        his.maxmedia = (uint32_t)prop;
        his.detnum = (uint32_t)c.sim.detnum;
        his.colcount = (uint32_t)reclen;
        his.totalphoton = (uint32_t)c.sim.nphoton;
        his.detected = (uint32_t)res.detected;
        his.savedphoton = (uint32_t)saved;
        his.unitinmm = c.unitinmm;
        his.normalizer = (float)normalizer;
        his.srcnum = 1;
        his.respin = 1;
        /* columns present: D(etid) S(cattering) P(ath) M(omentum) X(exit) V W(eight) bits */
        his.savedetflag = (c.sim.issaveexit ? 0x77 : 0x47) | (c.sim.ismomentum > 0 ? 0x08 : 0);
        his.seedbyte = c.issaveseed ? 16 : 0;
        snprintf(fname, sizeof(fname), "%s.mch", c.session);
        if (mmc_host_save_mch(fname, &his, res.detdata, saved)) die("can not write %s", fname);
        if (c.issaveseed) {
            /* the initial xorshift128+ state of each saved photon: the saved
             * state of a replayed photon, else the seeded hash of its id */
            FILE* fp = fopen(fname, "ab");
            if (!fp) die("can not append seeds to %s", fname);
            for (int64_t k = 0; k < saved; k++) {
                uint64_t st[2];
                int64_t id = res.detphoton[k];
                if (c.rseeds) {
                    st[0] = c.rseeds[2 * id];
                    st[1] = c.rseeds[2 * id + 1];
                } else {
                    mmc_hip_photon_seeds(c.sim.seed, id, 1, st);
                }
                if (fwrite(st, sizeof(uint64_t), 2, fp) != 2) die("can not append seeds to %s", fname);
            }
            fclose(fp);
        }
        printf("detected %lld photons\n", (long long)res.detected);
        if (c.sim.issaveexit == 2) {
            /* -x 2: time-resolved image of the area detector (mesh_getdetimage,
             * mesh_savedetimage: <session>.img, float [gate][y][x]) */
            const int xs = (int)c.detparam1[3], ys = (int)c.detparam2[3];
            if (xs < 1 || ys < 1) die("%s", "-x 2 needs the area detector's Param1/Param2 (pixel counts in w)");
            float* img = (float*)calloc((size_t)xs * ys * c.sim.maxgate, sizeof(float));
            if (!img || mmc_host_detimage(img, res.detdata, saved, reclen, prop, c.med, c.unitinmm, c.sim.tstep,
                                          c.sim.maxgate, c.detpos, c.detparam1, c.detparam2))
                die("%s", "can not form the detector image");
            snprintf(fname, sizeof(fname), "%s.img", c.session);
            FILE* fp = fopen(fname, "wb");
            if (!fp || fwrite(img, sizeof(float), (size_t)xs * ys * c.sim.maxgate, fp) != (size_t)xs * ys * c.sim.maxgate)
                die("can not write %s", fname);
            fclose(fp);
            free(img);
        }
    }
    if (c.savetraj) {
        /* the trajectory file: the history header with colcount 6 and
         * detected = 0 (src/mmc_host.c:452-459), named .mct by the filetag
         * rule of mesh_savedetphoton (src/mmc_mesh.c:1982) */
        int64_t kept = res.ntraj < c.maxjumpdebug ? res.ntraj : c.maxjumpdebug;
        if (res.ntraj > c.maxjumpdebug)
            printf("WARNING: %lld trajectory positions, more than --maxjumpdebug %lld; the rest were dropped\n",
                   (long long)res.ntraj, (long long)c.maxjumpdebug);
        mmc_mch_header his;
        memset(&his, 0, sizeof(his));
        memcpy(his.magic, "MCXH", 4);
        his.version = 1;
        his.maxmedia = (uint32_t)prop;
        his.detnum = (uint32_t)c.sim.detnum;
        his.colcount = MMC_TRAJ_RECLEN;
        his.totalphoton = (uint32_t)c.sim.nphoton;
        his.detected = 0;
        his.savedphoton = (uint32_t)kept;
        his.unitinmm = c.unitinmm;
        his.normalizer = (float)normalizer;
        his.srcnum = 1;
        his.respin = 1;
        snprintf(fname, sizeof(fname), "%s.mct", c.session);
        if (mmc_host_save_mch(fname, &his, res.traj, kept)) die("can not write %s", fname);
        printf("saved %lld trajectory positions\n", (long long)kept);
    }
    free(res.traj);
    free(field_im);
    free(field);
    free(res.detdata);
    free(res.detphoton);
    free(c.rseeds);
    free(c.rweight);
    free(c.rtime);
    free(elem);
    free(type);
    free(srcelem);
    free(evol);
    free(nvol);
    free(facenb);
    free(planes);
    free(c.node);
    free(c.elem5);
    free(c.med);
    free(c.pattern);
    return 0;
}
