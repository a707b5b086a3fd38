/*
 * mmc_output.c -- the volumetric output formats of `-F` for mesh data:
 * ascii (.dat), bin, nii, hdr/img (Analyze), mc2, tx3, jnii (JSON) and bnii
 * (binary JSON), following mesh_saveweight (src/mmc_mesh.c:1726-1800),
 * mcx_savedata (src/mmc_utils.c:919-993), mcx_savenii (src/mmc_utils.c:
 * 515-607), mcx_savejnii/mcx_savebnii (src/mmc_utils.c:622-905) and the JData
 * array annotation of mcx_jdataencode (src/mmc_utils.c:3071-3185).
 *
 * Mesh output is a [maxgate][datalen][srcnum] array of doubles (datalen =
 * nodes or elements).  NIfTI/Analyze describe it as dim = (srcnum, maxgate,
 * datalen, 1); JNIfTI as Dim = [datalen, maxgate, srcnum] with VoxelSize
 * [1, tstep, 1].  Dual-grid output is [maxgate][nz][ny][nx][srcnum]: NIfTI dim
 * = (nx, ny, nz, maxgate*srcnum) with the voxel edge as pixdim; JNIfTI Dim =
 * [nx, ny, nz, maxgate, srcnum], VoxelSize [step, step, step, tstep, 1].
 * JNIfTI data are zlib-compressed (the reference's default zipid) and, in the
 * text form, base64-encoded; `compress = 0` writes the raw array instead.
 * Binary JSON follows BJData draft 2 (little-endian, with the u/m/M unsigned
 * markers).
 */
#include <stddef.h>
#include <stdint.h>
#include <stdio.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <zlib.h>

#include "mmc_hip_host.h"

/* ------------------------------------------------------------ NIfTI-1 ---- */
/* The 348-byte NIfTI-1 header (public format, nifti.nimh.nih.gov). */
typedef struct {
    int32_t sizeof_hdr;
    char data_type[10];
    char db_name[18];
    int32_t extents;
    int16_t session_error;
    char regular;
    char dim_info;
    int16_t dim[8];
    float intent_p1, intent_p2, intent_p3;
    int16_t intent_code, datatype, bitpix, slice_start;
    float pixdim[8];
    float vox_offset, scl_slope, scl_inter;
    int16_t slice_end;
    char slice_code, xyzt_units;
    float cal_max, cal_min, slice_duration, toffset;
    int32_t glmax, glmin;
    char descrip[80];
    char aux_file[24];
    int16_t qform_code, sform_code;
    float quatern_b, quatern_c, quatern_d, qoffset_x, qoffset_y, qoffset_z;
    float srow_x[4], srow_y[4], srow_z[4];
    char intent_name[16];
    char magic[4];
} nii1_header;

_Static_assert(sizeof(nii1_header) == 348, "NIfTI-1 header must be 348 bytes");
_Static_assert(offsetof(nii1_header, dim) == 40, "dim offset");
_Static_assert(offsetof(nii1_header, datatype) == 70, "datatype offset");
_Static_assert(offsetof(nii1_header, pixdim) == 76, "pixdim offset");
_Static_assert(offsetof(nii1_header, vox_offset) == 108, "vox_offset offset");
_Static_assert(offsetof(nii1_header, xyzt_units) == 123, "xyzt_units offset");
_Static_assert(offsetof(nii1_header, magic) == 344, "magic offset");

enum { NII_FLOAT64 = 64, NII_UNITS_MM = 2, NII_UNITS_USEC = 24 };

static int write_all(FILE* fp, const void* p, size_t n) { return fwrite(p, 1, n, fp) == n ? 0 : -1; }

/* shape of an output array: ndim dims (fastest first) with voxel sizes, and
 * the 4 NIfTI dims / 3 spatial pixdims */
typedef struct {
    int ndim;
    uint32_t dims[5];
    float vox[5];
    int32_t nii[4];
    float pix;
} out_shape;

static int16_t nii_dim(int32_t v) { return (int16_t)(v > 32767 ? 32767 : v); }

static int save_nifti(const char* session, const double* f, int64_t n, const out_shape* sh, float tstep,
                      int analyze) {
    nii1_header h;
    memset(&h, 0, sizeof(h));
    h.sizeof_hdr = 348;
    h.dim[0] = 4;
    for (int k = 0; k < 4; k++) h.dim[1 + k] = nii_dim(sh->nii[k]);
    h.datatype = NII_FLOAT64;
    h.bitpix = 64;
    h.pixdim[1] = h.pixdim[2] = h.pixdim[3] = sh->pix;
    h.pixdim[4] = tstep * 1e6f;
    h.xyzt_units = NII_UNITS_MM | NII_UNITS_USEC;
    memcpy(h.magic, analyze ? "ni1\0" : "n+1\0", 4);
    h.vox_offset = analyze ? 0.f : 352.f;
    char fname[4096];
    snprintf(fname, sizeof(fname), "%s.%s", session, analyze ? "hdr" : "nii");
    FILE* fp = fopen(fname, "wb");
    if (!fp) return -1;
    int rc = write_all(fp, &h, sizeof(h));
    if (!analyze) {
        const uint8_t ext[4] = {0, 0, 0, 0};
        if (!rc) rc = write_all(fp, ext, 4);
        if (!rc) rc = write_all(fp, f, sizeof(double) * (size_t)n);
        fclose(fp);
        return rc;
    }
    fclose(fp);
    if (rc) return rc;
    snprintf(fname, sizeof(fname), "%s.img", session);
    if (!(fp = fopen(fname, "wb"))) return -1;
    rc = write_all(fp, f, sizeof(double) * (size_t)n);
    fclose(fp);
    return rc;
}

/* ------------------------------------------------------------- base64 ---- */
static char* base64(const uint8_t* in, size_t n, size_t* outlen) {
    static const char tab[] = "ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789+/";
    size_t m = 4 * ((n + 2) / 3);
    char* out = (char*)malloc(m + 1);
    if (!out) return NULL;
    size_t j = 0;
    for (size_t i = 0; i < n; i += 3) {
        uint32_t v = (uint32_t)in[i] << 16;
        if (i + 1 < n) v |= (uint32_t)in[i + 1] << 8;
        if (i + 2 < n) v |= in[i + 2];
        out[j++] = tab[(v >> 18) & 63];
        out[j++] = tab[(v >> 12) & 63];
        out[j++] = i + 1 < n ? tab[(v >> 6) & 63] : '=';
        out[j++] = i + 2 < n ? tab[v & 63] : '=';
    }
    out[j] = 0;
    if (outlen) *outlen = j;
    return out;
}

static uint8_t* zlib_pack(const void* in, size_t n, size_t* outlen) {
    uLongf cap = compressBound((uLong)n);
    uint8_t* out = (uint8_t*)malloc(cap);
    if (!out) return NULL;
    if (compress2(out, &cap, (const Bytef*)in, (uLong)n, Z_DEFAULT_COMPRESSION) != Z_OK) {
        free(out);
        return NULL;
    }
    *outlen = cap;
    return out;
}

static const char* output_description(int32_t outputtype) {
    switch (outputtype) {
        case MMC_OUT_FLUX: return "MMC volumetric output: Fluence rate (W/mm^2)";
        case MMC_OUT_FLUENCE: return "MMC volumetric output: Fluence (J/mm^2)";
        case MMC_OUT_ENERGY: return "MMC volumetric output: Voxel-wise energy deposit (J)";
        default: return "MMC volumetric output";
    }
}

/* ------------------------------------------------------- JNIfTI (text) ---- */
static int save_jnii(const char* session, const double* f, int64_t n, const out_shape* sh, int32_t maxgate,
                     int32_t outputtype, int compress) {
    const int ndim = sh->ndim;
    size_t rawbytes = sizeof(double) * (size_t)n, zbytes = 0, blen = 0;
    uint8_t* z = NULL;
    if (compress && !(z = zlib_pack(f, rawbytes, &zbytes))) return -1;
    char* b64 = base64(compress ? z : (const uint8_t*)f, compress ? zbytes : rawbytes, &blen);
    free(z);
    if (!b64) return -1;
    char fname[4096];
    snprintf(fname, sizeof(fname), "%s.jnii", session);
    FILE* fp = fopen(fname, "wt");
    if (!fp) {
        free(b64);
        return -1;
    }
    fprintf(fp, "{\n\t\"_DataInfo_\": {\n\t\t\"JNIFTIVersion\": \"0.5\",\n"
                "\t\t\"Comment\": \"Created by the MI355X mesh Monte Carlo tracer\",\n"
                "\t\t\"AnnotationFormat\": \"https://neurojson.org/jnifti/draft1\",\n"
                "\t\t\"SerialFormat\": \"https://json.org\"\n\t},\n");
    fprintf(fp, "\t\"NIFTIHeader\": {\n\t\t\"NIIHeaderSize\": 348,\n\t\t\"Dim\": [");
    for (int i = 0; i < ndim; i++) fprintf(fp, "%s%u", i ? ", " : "", sh->dims[i]);
    fprintf(fp, "],\n\t\t\"Param1\": 0,\n\t\t\"Param2\": 0,\n\t\t\"Param3\": 0,\n\t\t\"Intent\": 0,\n"
                "\t\t\"DataType\": \"double\",\n\t\t\"BitDepth\": 64,\n\t\t\"FirstSliceID\": 0,\n"
                "\t\t\"VoxelSize\": [");
    for (int i = 0; i < ndim; i++) fprintf(fp, "%s%.9g", i ? ", " : "", (double)sh->vox[i]);
    fprintf(fp, "],\n\t\t\"Orientation\": {\"x\": \"r\", \"y\": \"a\", \"z\": \"s\"},\n"
                "\t\t\"ScaleSlope\": 1,\n\t\t\"ScaleOffset\": 0,\n\t\t\"LastSliceID\": %d,\n"
                "\t\t\"SliceType\": 1,\n\t\t\"Unit\": {\"L\": \"mm\", \"T\": \"s\"},\n"
                "\t\t\"MaxIntensity\": 1,\n\t\t\"MinIntensity\": 0,\n\t\t\"SliceTime\": 0,\n"
                "\t\t\"TimeOffset\": 0,\n\t\t\"Description\": \"%s\",\n\t\t\"AuxFile\": \"\",\n"
                "\t\t\"QForm\": 0,\n\t\t\"SForm\": 1,\n\t\t\"Quatern\": {\"b\": 0, \"c\": 0, \"d\": 0},\n"
                "\t\t\"QuaternOffset\": {\"x\": 0, \"y\": 0, \"z\": 0},\n"
                "\t\t\"Affine\": [[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 1, 0]],\n"
                "\t\t\"Name\": \"%s\",\n\t\t\"NIIFormat\": \"jnifti\"\n\t},\n",
            maxgate, output_description(outputtype), session);
    fprintf(fp, "\t\"NIFTIData\": {\n\t\t\"_ArrayType_\": \"double\",\n\t\t\"_ArraySize_\": [");
    for (int i = 0; i < ndim; i++) fprintf(fp, "%s%u", i ? ", " : "", sh->dims[i]);
    fprintf(fp, "],\n");
    if (compress)
        fprintf(fp, "\t\t\"_ArrayZipType_\": \"zlib\",\n\t\t\"_ArrayZipSize_\": %lld,\n\t\t\"_ArrayZipData_\": \"",
                (long long)n);
    else
        fprintf(fp, "\t\t\"_ArrayData_\": \"");
    int rc = write_all(fp, b64, blen);
    fprintf(fp, "\"\n\t}\n}\n");
    free(b64);
    if (fclose(fp)) rc = -1;
    return rc;
}

/* ---------------------------------------------- binary JSON (BJData 2) ---- */
typedef struct {
    uint8_t* p;
    size_t n, cap;
    int err;
} bj_buf;

static void bj_put(bj_buf* b, const void* src, size_t n) {
    if (b->err) return;
    if (b->n + n > b->cap) {
        size_t cap = (b->cap ? b->cap * 2 : 4096);
        while (cap < b->n + n) cap *= 2;
        uint8_t* q = (uint8_t*)realloc(b->p, cap);
        if (!q) {
            b->err = 1;
            return;
        }
        b->p = q;
        b->cap = cap;
    }
    memcpy(b->p + b->n, src, n);
    b->n += n;
}
static void bj_c(bj_buf* b, char c) { bj_put(b, &c, 1); }
/* smallest non-negative integer encoding (lengths and counts) */
static void bj_len(bj_buf* b, uint64_t v) {
    if (v < 256) {
        uint8_t x = (uint8_t)v;
        bj_c(b, 'U');
        bj_put(b, &x, 1);
    } else if (v < 65536) {
        uint16_t x = (uint16_t)v;
        bj_c(b, 'u');
        bj_put(b, &x, 2);
    } else if (v < 4294967296ull) {
        uint32_t x = (uint32_t)v;
        bj_c(b, 'm');
        bj_put(b, &x, 4);
    } else {
        bj_c(b, 'M');
        bj_put(b, &v, 8);
    }
}
static void bj_key(bj_buf* b, const char* k) {
    size_t n = strlen(k);
    bj_len(b, n);
    bj_put(b, k, n);
}
static void bj_str(bj_buf* b, const char* s) {
    size_t n = strlen(s);
    bj_c(b, 'S');
    bj_len(b, n);
    bj_put(b, s, n);
}
static void bj_kstr(bj_buf* b, const char* k, const char* s) { bj_key(b, k); bj_str(b, s); }
static void bj_ku8(bj_buf* b, const char* k, uint8_t v) { bj_key(b, k); bj_c(b, 'U'); bj_put(b, &v, 1); }
static void bj_ku16(bj_buf* b, const char* k, uint16_t v) { bj_key(b, k); bj_c(b, 'u'); bj_put(b, &v, 2); }
static void bj_ku32(bj_buf* b, const char* k, uint32_t v) { bj_key(b, k); bj_c(b, 'm'); bj_put(b, &v, 4); }
static void bj_kchar(bj_buf* b, const char* k, char v) { bj_key(b, k); bj_c(b, 'C'); bj_c(b, v); }
/* optimized typed array: [ $ type # count raw */
static void bj_typed(bj_buf* b, char type, const void* data, size_t count, size_t elem) {
    bj_c(b, '[');
    bj_c(b, '$');
    bj_c(b, type);
    bj_c(b, '#');
    bj_len(b, count);
    bj_put(b, data, count * elem);
}

static int save_bnii(const char* session, const double* f, int64_t n, const out_shape* sh, int32_t maxgate,
                     int32_t outputtype, int compress) {
    const int ndim = sh->ndim;
    const uint32_t* dims = sh->dims;
    const float* vox = sh->vox;
    const int32_t affine[3][4] = {{1, 0, 0, 0}, {0, 1, 0, 0}, {0, 0, 1, 0}};
    bj_buf b = {NULL, 0, 0, 0};
    bj_c(&b, '{');
    bj_key(&b, "_DataInfo_");
    bj_c(&b, '{');
    bj_kstr(&b, "JNIFTIVersion", "0.5");
    bj_kstr(&b, "Comment", "Created by the MI355X mesh Monte Carlo tracer");
    bj_kstr(&b, "AnnotationFormat", "https://neurojson.org/jnifti/draft1");
    bj_kstr(&b, "SerialFormat", "https://neurojson.org/bjdata/draft2");
    bj_c(&b, '}');
    bj_key(&b, "NIFTIHeader");
    bj_c(&b, '{');
    bj_ku16(&b, "NIIHeaderSize", 348);
    bj_key(&b, "Dim");
    bj_typed(&b, 'm', dims, (size_t)ndim, 4);
    bj_ku8(&b, "Param1", 0);
    bj_ku8(&b, "Param2", 0);
    bj_ku8(&b, "Param3", 0);
    bj_ku8(&b, "Intent", 0);
    bj_kstr(&b, "DataType", "double");
    bj_ku8(&b, "BitDepth", 64);
    bj_ku8(&b, "FirstSliceID", 0);
    bj_key(&b, "VoxelSize");
    bj_typed(&b, 'd', vox, (size_t)ndim, 4);
    bj_key(&b, "Orientation");
    bj_c(&b, '{');
    bj_kchar(&b, "x", 'r');
    bj_kchar(&b, "y", 'a');
    bj_kchar(&b, "z", 's');
    bj_c(&b, '}');
    bj_ku8(&b, "ScaleSlope", 1);
    bj_ku8(&b, "ScaleOffset", 0);
    bj_ku32(&b, "LastSliceID", (uint32_t)maxgate);
    bj_ku8(&b, "SliceType", 1);
    bj_key(&b, "Unit");
    bj_c(&b, '{');
    bj_kstr(&b, "L", "mm");
    bj_kstr(&b, "T", "s");
    bj_c(&b, '}');
    bj_ku32(&b, "MaxIntensity", 1);
    bj_ku32(&b, "MinIntensity", 0);
    bj_ku8(&b, "SliceTime", 0);
    bj_ku8(&b, "TimeOffset", 0);
    bj_kstr(&b, "Description", output_description(outputtype));
    bj_kstr(&b, "AuxFile", "");
    bj_ku8(&b, "QForm", 0);
    bj_ku8(&b, "SForm", 1);
    bj_key(&b, "Quatern");
    bj_c(&b, '{');
    bj_ku8(&b, "b", 0);
    bj_ku8(&b, "c", 0);
    bj_ku8(&b, "d", 0);
    bj_c(&b, '}');
    bj_key(&b, "QuaternOffset");
    bj_c(&b, '{');
    bj_ku8(&b, "x", 0);
    bj_ku8(&b, "y", 0);
    bj_ku8(&b, "z", 0);
    bj_c(&b, '}');
    bj_key(&b, "Affine");
    bj_c(&b, '[');
    for (int r = 0; r < 3; r++) bj_typed(&b, 'l', affine[r], 4, 4);
    bj_c(&b, ']');
    bj_kstr(&b, "Name", session);
    bj_kstr(&b, "NIIFormat", "jnifti");
    bj_c(&b, '}');
    bj_key(&b, "NIFTIData");
    bj_c(&b, '{');
    bj_kstr(&b, "_ArrayType_", "double");
    bj_key(&b, "_ArraySize_");
    bj_typed(&b, 'm', dims, (size_t)ndim, 4);
    if (compress) {
        size_t zbytes = 0;
        uint8_t* z = zlib_pack(f, sizeof(double) * (size_t)n, &zbytes);
        if (!z) {
            free(b.p);
            return -1;
        }
        bj_kstr(&b, "_ArrayZipType_", "zlib");
        bj_ku32(&b, "_ArrayZipSize_", (uint32_t)n);
        bj_key(&b, "_ArrayZipData_");
        bj_typed(&b, 'U', z, zbytes, 1);
        free(z);
    } else {
        bj_key(&b, "_ArrayData_");
        bj_typed(&b, 'D', f, (size_t)n, 8);
    }
    bj_c(&b, '}');
    bj_c(&b, '}');
    int rc = b.err ? -1 : 0;
    if (!rc) {
        char fname[4096];
        snprintf(fname, sizeof(fname), "%s.bnii", session);
        FILE* fp = fopen(fname, "wb");
        rc = fp ? write_all(fp, b.p, b.n) : -1;
        if (fp && fclose(fp)) rc = -1;
    }
    free(b.p);
    return rc;
}

/* ------------------------------------------------------- dispatcher ----- */
/* the .dat text of mesh_saveweight (src/mmc_mesh.c:1760-1797): "%d\t%e" per
 * value, or "%d\t%d\t%e" (value, pattern) with several patterns */
static int save_text(const char* session, const double* f, int32_t datalen, int32_t maxgate, int32_t srcnum) {
    char fname[4096];
    snprintf(fname, sizeof(fname), "%s.dat", session);
    if (srcnum == 1) return mmc_host_save_dat(fname, f, datalen, maxgate);
    FILE* fp = fopen(fname, "wt");
    if (!fp) return -1;
    int rc = 0;
    for (int32_t g = 0; g < maxgate && !rc; g++)
        for (int32_t j = 0; j < datalen && !rc; j++)
            for (int32_t k = 0; k < srcnum && !rc; k++)
                if (fprintf(fp, "%d\t%d\t%e\n", j + 1, k + 1, f[((size_t)g * datalen + j) * srcnum + k]) <= 0)
                    rc = -2;
    if (fclose(fp)) rc = -1;
    return rc;
}

static int save_shaped(const char* session, int32_t format, const double* field, int32_t datalen, int32_t maxgate,
                       int32_t srcnum, float tstep, int32_t outputtype, int32_t compress, const out_shape* sh,
                       const int32_t tx3dim[3]) {
    const int64_t n = (int64_t)datalen * maxgate * srcnum;
    char fname[4096];
    switch (format) {
        case MMC_FMT_ASCII:
            return save_text(session, field, datalen, maxgate, srcnum);
        case MMC_FMT_BIN:
        case MMC_FMT_MC2: {
            snprintf(fname, sizeof(fname), "%s.%s", session, format == MMC_FMT_BIN ? "bin" : "mc2");
            return mmc_host_save_bin(fname, field, n);
        }
        case MMC_FMT_NII:
        case MMC_FMT_HDR:
            return save_nifti(session, field, n, sh, tstep, format == MMC_FMT_HDR);
        case MMC_FMT_TX3: {
            snprintf(fname, sizeof(fname), "%s.tx3", session);
            FILE* fp = fopen(fname, "wb");
            if (!fp) return -1;
            const uint32_t glformat = 0x8814; /* GL_RGBA32F */
            int rc = write_all(fp, &glformat, 4);
            if (!rc) rc = write_all(fp, tx3dim, 3 * sizeof(int32_t));
            if (!rc) rc = write_all(fp, field, sizeof(double) * (size_t)n);
            if (fclose(fp)) rc = -1;
            return rc;
        }
        case MMC_FMT_JNII:
            return save_jnii(session, field, n, sh, maxgate, outputtype, compress);
        case MMC_FMT_BNII:
            return save_bnii(session, field, n, sh, maxgate, outputtype, compress);
        default:
            return -1;
    }
}

/* mesh data: the reference describes it as cfg->dim = (srcnum, maxgate,
 * datalen) (src/mmc_mesh.c:1746-1756) and JNIfTI dims [datalen, maxgate,
 * srcnum] (src/mmc_utils.c:941-956) */
int32_t mmc_host_save_field(const char* session, int32_t format, const double* field, int32_t datalen,
                            int32_t maxgate, int32_t srcnum, float tstep, int32_t outputtype, int32_t compress) {
    if (!session || !field || datalen < 1 || maxgate < 1 || srcnum < 1) return -1;
    out_shape sh;
    memset(&sh, 0, sizeof(sh));
    sh.ndim = 3;
    sh.dims[0] = (uint32_t)datalen;
    sh.dims[1] = (uint32_t)maxgate;
    sh.dims[2] = (uint32_t)srcnum;
    sh.vox[0] = 1.f;
    sh.vox[1] = tstep;
    sh.vox[2] = 1.f;
    sh.nii[0] = srcnum;
    sh.nii[1] = maxgate;
    sh.nii[2] = datalen;
    sh.nii[3] = 1;
    sh.pix = 1.f;
    const int32_t tx3dim[3] = {srcnum, maxgate, datalen};
    return save_shaped(session, format, field, datalen, maxgate, srcnum, tstep, outputtype, compress, &sh, tx3dim);
}

/* dual-grid data: cfg->dim voxels of edge cfg->steps (src/mmc_utils.c:
 * 515-545, 937-963) */
int32_t mmc_host_save_grid(const char* session, int32_t format, const double* field, const int32_t dim[3],
                           float step, int32_t maxgate, int32_t srcnum, float tstep, int32_t outputtype,
                           int32_t compress) {
    if (!session || !field || !dim || dim[0] < 1 || dim[1] < 1 || dim[2] < 1 || maxgate < 1 || srcnum < 1)
        return -1;
    const int64_t ncell = (int64_t)dim[0] * dim[1] * dim[2];
    if (ncell >= (1ll << 31)) return -1;
    out_shape sh;
    memset(&sh, 0, sizeof(sh));
    sh.ndim = 5;
    for (int k = 0; k < 3; k++) {
        sh.dims[k] = (uint32_t)dim[k];
        sh.vox[k] = step;
        sh.nii[k] = dim[k];
    }
    sh.dims[3] = (uint32_t)maxgate;
    sh.dims[4] = (uint32_t)srcnum;
    sh.vox[3] = tstep;
    sh.vox[4] = 1.f;
    sh.nii[3] = maxgate * srcnum;
    sh.pix = step;
    return save_shaped(session, format, field, (int32_t)ncell, maxgate, srcnum, tstep, outputtype, compress, &sh,
                       dim);
}

int32_t mmc_host_format_id(const char* name) {
    static const char* names[] = {"ascii", "bin", "nii", "hdr", "mc2", "tx3", "jnii", "bnii"};
    if (!name) return -1;
    for (int i = 0; i < (int)(sizeof(names) / sizeof(names[0])); i++)
        if (!strcmp(name, names[i])) return i;
    return -1;
}

/* ---- area detectors (post-processing of detected records) ---- */

/* record columns: detid, nscat[prop], ppath[prop], (mom[prop]), p[3], v[3], w0 */
static int det_bin(const float* row, int32_t colcount, const float detpos0[2], const float p1[4], const float p2[4],
                   int* xi, int* yi) {
    const float xrange = p1[0] + p2[0], yrange = p1[1] + p2[1];
    const int xsize = (int)p1[3], ysize = (int)p2[3];
    const float xloc = row[colcount - 7], yloc = row[colcount - 6];
    *xi = (int)((xloc - detpos0[0]) / xrange * xsize);
    *yi = (int)((yloc - detpos0[1]) / yrange * ysize);
    return !(*xi < 0 || *xi > xsize - 1 || *yi < 0 || *yi > ysize - 1);
}

int32_t mmc_host_detimage(float* detmap, const float* rec, int64_t count, int32_t colcount, int32_t prop,
                          const float* med, float unitinmm, float tstep, int32_t maxgate, const float detpos0[2],
                          const float detparam1[4], const float detparam2[4]) {
    if (!detmap || (!rec && count > 0) || !med || !detpos0 || !detparam1 || !detparam2 || prop < 1 ||
        colcount < 2 * prop + 8 || maxgate < 1 || !(tstep > 0.f) || detparam1[3] < 1.f || detparam2[3] < 1.f)
        return -1;
    const float R_C0 = 3.335640951981520e-12f; /* src/mmc_const.h:45 */
    const int xsize = (int)detparam1[3], ysize = (int)detparam2[3];
    for (int64_t i = 0; i < count; i++) {
        const float* row = rec + (size_t)i * colcount;
        float path = 0.f, weight = row[colcount - 1];
        for (int j = 1; j <= prop; j++) {
            path += row[j + prop] * med[4 * j + 3];
            weight *= expf(-row[j + prop] * med[4 * j] * unitinmm);
        }
        /* the reference truncates the optical path to an integer before
         * scaling it to a gate (src/mmc_mesh.c:2053) */
        int ntg = (int)((float)(int)path * R_C0 / tstep);
        if (ntg > maxgate - 1) ntg = maxgate - 1;
        int xi, yi;
        if (!det_bin(row, colcount, detpos0, detparam1, detparam2, &xi, &yi)) continue;
        detmap[(size_t)ntg * xsize * ysize + (size_t)yi * xsize + xi] += weight;
    }
    return 0;
}

int32_t mmc_host_detweight(float* weight, const float* rec, int64_t count, int32_t colcount, const float detpos0[2],
                           const float detparam1[4], const float detparam2[4], const float* detpattern) {
    if (!weight || (!rec && count > 0) || !detpos0 || !detparam1 || !detparam2 || !detpattern ||
        detparam1[3] < 1.f || detparam2[3] < 1.f || colcount < 8)
        return -1;
    const int xsize = (int)detparam1[3];
    for (int64_t i = 0; i < count; i++) {
        int xi, yi;
        if (!det_bin(rec + (size_t)i * colcount, colcount, detpos0, detparam1, detparam2, &xi, &yi))
            return -2; /* "photon location not within the detection plane" */
        weight[i] = detpattern[yi * xsize + xi];
    }
    return 0;
}
