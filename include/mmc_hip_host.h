/*
 * mmc_hip_host.h -- host-side mesh preparation, normalization and output writers
 * of the MI355X photon tracer (part of libmmc_hip.so; no GPU calls).
 *
 * These restate the reference's mmc_prep() chain (src/mmc_host.c:136-165):
 * mesh_getvolume (src/mmc_mesh.c:910-950), mesh_getfacenb
 * (src/mmc_highorder.cpp:128-159), tracer_build (src/mmc_mesh.c:1496-1602),
 * tracer_prep (src/mmc_mesh.c:1241-1483), mesh_initelem / mesh_barycentric
 * (src/mmc_mesh.c:1060-1206) and mesh_normalize (src/mmc_mesh.c:2154-2279),
 * plus the writers mesh_saveweight (src/mmc_mesh.c:1726-1800), mcx_savedata
 * (src/mmc_utils.c:919-993) and mcx_savedetphoton (src/mmc_utils.c:4309-4342).
 */
#ifndef MMC_HIP_HOST_H
#define MMC_HIP_HOST_H

#include <stdint.h>
#include "mmc_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* .mch header, field-compatible with `history` (src/mmc_utils.h:168-184) */
typedef struct mmc_mch_header {
    char magic[4];
    uint32_t version, maxmedia, detnum, colcount, totalphoton, detected, savedphoton;
    float unitinmm;
    uint32_t seedbyte;
    float normalizer;
    uint32_t srcnum;
    int32_t respin;
    uint32_t savedetflag;
    int32_t reserved[2];
} mmc_mch_header;

/* Element volumes, node volumes and orientation fix (swaps nodes 3,4 of
 * negatively oriented elements in place). */
void mmc_host_volume(int32_t nn, const float* node, int32_t ne, int32_t* elem,
                     const int32_t* type, float* evol, float* nvol);
/* Face-neighbour table: facenb[i*4+j] = element sharing face j of element i
 * (faces {0,1,2},{0,1,3},{0,2,3},{1,2,3}), 0 on the exterior. */
int32_t mmc_host_facenb(int32_t nn, int32_t ne, const int32_t* elem, int32_t* facenb);
/* Face planes: per element {nx[4],ny[4],nz[4],d[4]}, unit outward normals. */
void mmc_host_planes(int32_t ne, const float* node, const int32_t* elem, float* planes);
/* Barycentric test of srcpos in element e0 (1-based): 0 and bary if inside. */
int32_t mmc_host_barycentric(const float* node, const int32_t* elem, int32_t e0,
                             const float* srcpos, float* bary);
/* First element (index order) enclosing srcpos, 0 if none. */
int32_t mmc_host_initelem(int32_t ne, const float* node, const int32_t* elem,
                          const float* srcpos, float* bary);
/* Effective reflection coefficient (mesh_getreff). */
double mmc_host_reff(double n_in, double n_out);
/* Surface-node volume correction for nodal output; facenb must still hold 0
 * on exterior faces. */
void mmc_host_surface_nvol(int32_t nn, int32_t ne, const int32_t* elem, const int32_t* type,
                           const int32_t* facenb, const float* med, int32_t prop,
                           int32_t isreflect, float* nvol);
/* Relabels wide-field detector elements (label -2) as prop+1, as
 * mesh_srcdetelem + mesh_loadmedia do (src/mmc_mesh.c:390-427, 511-527);
 * returns their count (mmc_hip_mesh.isextdet = count > 0). */
int32_t mmc_host_detelem(int32_t ne, int32_t* type, int32_t prop);
/* Nodal output of the reference GPU backend under the Badouel tracer
 * (src/mmc_cl_host.c:973-1024): a quarter of each element value of every
 * gate and pattern to each of the element's nodes; felem [gate][elem][pattern]
 * (an element-output run), fnode [gate][node][pattern] (added to).  The
 * library's own Badouel nodal output (basisorder 1) follows the reference's
 * CPU path instead: a third of each step's deposit to each node of the exit
 * face.  0, or -1 for bad arguments. */
int32_t mmc_host_elem_to_nodes(int32_t ne, int32_t nn, const int32_t* elem, int32_t maxgate, int32_t srcnum,
                               const double* felem, double* fnode);
/* Source elements (label -1): writes their 1-based ids in index order to
 * srcelem (if not NULL; ne entries suffice), relabels them 0 and returns their
 * count -- the candidate list of area sources' launch search
 * (mesh_srcdetelem, src/mmc_mesh.c:390-427). */
int32_t mmc_host_srcelem(int32_t ne, int32_t* type, int32_t* srcelem);
/* Numbers exterior faces -1,-2,... in element-major order; returns the count. */
int32_t mmc_host_number_exterior(int32_t ne, int32_t* facenb);
/* Dual-grid extent (mesh_createdualmesh, src/mmc_mesh.c:349-381): the
 * node bounding box widened by 1e-6 gives the grid origin nmin and dim =
 * (int)(extent / step) + 1 voxels per axis.  Returns 0, or -1 for a bad
 * step. */
int32_t mmc_host_grid(const float* node, int32_t nn, float step, float nmin[3], int32_t dim[3]);
/* Per element the products of the linear basis gradients times the element
 * volume, <grad phi_i . grad phi_j> * evol, upper triangle packed as
 * [00,01,02,03,11,12,13,22,23,33] into ddd[ne*10] (mesh_deldotdel,
 * src/mmc_mesh.c:990-1046).  0, or -1 for a node id outside 1..nn. */
int32_t mmc_host_deldotdel(const float* node, int32_t nn, const int32_t* elem, int32_t ne, const float* evol,
                           double* ddd);
/* Area detector over the plane at detpos[0].xy spanned by detparam1.xy and
 * detparam2.xy, detparam1.w x detparam2.w pixels.  mmc_host_detimage (-x 2;
 * mesh_getdetimage, src/mmc_mesh.c:2029-2078) adds each detected record's
 * weight w0 * prod exp(-mua_j L_j unitinmm) to detmap[gate][y][x] (float,
 * maxgate x ysize x xsize, zeroed by the caller), the gate from the optical
 * path sum n_j L_j truncated to an integer as the reference does; records
 * outside the plane are skipped.  rec rows: detid, nscat[prop], ppath[prop],
 * (mom[prop]), exit p[3], v[3], w0 (colcount columns).  mmc_host_detweight
 * (mesh_getdetweight, src/mmc_mesh.c:2121-2138) gives each record the
 * detector pattern's value at its exit pixel (the replay weight of a
 * patterned detector); -2 if a record lies outside the plane. */
int32_t mmc_host_detimage(float* detmap, const float* rec, int64_t count, int32_t colcount, int32_t prop,
                          const float* med, float unitinmm, float tstep, int32_t maxgate, const float detpos0[2],
                          const float detparam1[4], const float detparam2[4]);
int32_t mmc_host_detweight(float* weight, const float* rec, int64_t count, int32_t colcount, const float detpos0[2],
                           const float detparam1[4], const float detparam2[4], const float* detpattern);
/* Normalizes a field in place and returns the factor (mesh_normalize). */
double mmc_host_normalize(const mmc_hip_sim* sim, int32_t nn, int32_t ne,
                          const int32_t* elem, const int32_t* type, const float* med,
                          const float* evol, const float* nvol, double* field,
                          double absorbweight, double launchweight);
/* The same for pattern slot `pair` of a field shared by srcnum patterns
 * (values strided by srcnum; mesh_normalize's pair argument,
 * src/mmc_mesh.c:2154), with that pattern's absorbed and launched energy. */
double mmc_host_normalize_pair(const mmc_hip_sim* sim, int32_t nn, int32_t ne,
                               const int32_t* elem, const int32_t* type, const float* med,
                               const float* evol, const float* nvol, double* field,
                               double absorbweight, double launchweight, int32_t pair);
/* The same for a dual-grid field of ncells voxels per gate (fluence per
 * unit launched energy and unitinmm^3, src/mmc_mesh.c:2205-2206). */
double mmc_host_normalize_grid(const mmc_hip_sim* sim, int64_t ncells, double* field, double absorbweight,
                               double launchweight, int32_t pair);

/* Writers.  Return 0 on success. */
int32_t mmc_host_save_dat(const char* fname, const double* field, int32_t datalen,
                          int32_t maxgate);
int32_t mmc_host_save_bin(const char* fname, const double* field, int64_t count);
int32_t mmc_host_save_mch(const char* fname, const mmc_mch_header* his, const float* data,
                          int64_t count);

/* Output formats of -F (enum TOutputFormat order, src/mmc_utils.h:112). */
#define MMC_FMT_ASCII 0   /* <session>.dat, "%d\t%e" per value */
#define MMC_FMT_BIN   1   /* <session>.bin, raw doubles */
#define MMC_FMT_NII   2   /* <session>.nii, NIfTI-1, dim = (srcnum, maxgate, datalen) */
#define MMC_FMT_HDR   3   /* <session>.hdr + .img, Analyze 7.5 */
#define MMC_FMT_MC2   4   /* <session>.mc2, raw doubles */
#define MMC_FMT_TX3   5   /* <session>.tx3, GL_RGBA32F + int dim[3] + doubles */
#define MMC_FMT_JNII  6   /* <session>.jnii, JNIfTI JSON (zlib + base64, or raw base64) */
#define MMC_FMT_BNII  7   /* <session>.bnii, JNIfTI binary JSON (BJData draft 2) */
/* Saves a [maxgate][datalen][srcnum] double field as <session>.<ext>
 * (mesh_saveweight + mcx_savedata, src/mmc_mesh.c:1726-1800,
 * src/mmc_utils.c:919-993); compress selects zlib for jnii/bnii. */
int32_t mmc_host_save_field(const char* session, int32_t format, const double* field, int32_t datalen,
                            int32_t maxgate, int32_t srcnum, float tstep, int32_t outputtype,
                            int32_t compress);
/* Saves a dual-grid field ([maxgate][nz][ny][nx][srcnum] doubles over dim
 * voxels of edge step) in the same formats (mcx_savedata with cfg->dim,
 * src/mmc_utils.c:919-993). */
int32_t mmc_host_save_grid(const char* session, int32_t format, const double* field, const int32_t dim[3],
                           float step, int32_t maxgate, int32_t srcnum, float tstep, int32_t outputtype,
                           int32_t compress);
/* Index of a -F name ("ascii", "bin", "nii", "hdr", "mc2", "tx3", "jnii",
 * "bnii"), or -1. */
int32_t mmc_host_format_id(const char* name);

/* ---- agreement with the host C library's float functions ----------------
 * The reference CPU tracer takes every scattering length, scattering azimuth
 * and isotropic zenith from glibc's logf / sincosf / acosf
 * (src/mmc_rand_common.h:50-51, src/mmc_mesh.c:1662-1690,
 * src/mmc_raytrace.c:2434-2480), which are within one ulp but not correctly
 * rounded.  Each such call sees a function of one generator output
 * u = k / 2^23 (k: the 23 mantissa bits of xorshift128+,
 * src/mmc_rand_xorshift128p.c:55-70).  The GPU evaluates each value in double
 * (mmc_log_d, mmc_sincos_d, mmc_acos_d of mmc_fp32math.h) and rounds it to
 * float f; a table over k holds one bit per value, set when the C library
 * returns instead the float o next to f on the far side of the double, and
 * per region of 2^MMC_LIBM_REGSHIFT consecutive k and per value a threshold D
 * (float, -1 for none): every set bit of the region has its double within
 * D |o - f| of the midpoint (f + o)/2, so the GPU reads bits only there.
 *   which         values of generator output k            bits per k
 *   MMC_FIX_LOG   logf(u + 1e-6f)                          1
 *   MMC_FIX_AZ    sincosf((float)(2 pi u))                 2 (sin, cos)
 *   MMC_FIX_ISO   t = acosf(2u - 1); sincosf(t)            4 (acos, sin, cos, unused)
 *   MMC_FIX_PI    sincosf((float)(pi u))                   2 (sin, cos)
 * Bit b of value v at k: byte (k * bits) / 8, bit (k * bits) % 8 + v.
 * mmc_host_libm_table fills the bits (mmc_host_libm_table_bytes bytes) and
 * the thresholds (MMC_LIBM_NREG * mmc_host_libm_nval floats, [region][value])
 * from the C library this process runs with; returns the number of values the
 * rule cannot reproduce (the C library's result is neither f nor o; 0 on
 * glibc), or -1 on bad arguments.  Needs a host CPU with FMA. */
#define MMC_FIX_LOG 0
#define MMC_FIX_AZ  1
#define MMC_FIX_ISO 2
#define MMC_FIX_PI  3
#define MMC_FIXB_LOG (1 << 20)
#define MMC_FIXB_AZ  (1 << 21)
#define MMC_FIXB_ISO (1 << 22)
#define MMC_FIXB_PI  (1 << 21)
#define MMC_LIBM_REGSHIFT 16
#define MMC_LIBM_NREG (1 << (23 - MMC_LIBM_REGSHIFT))
int32_t mmc_host_libm_nval(int32_t which);
int64_t mmc_host_libm_table_bytes(int32_t which);
int32_t mmc_host_libm_table(int32_t which, uint8_t* bits, float* thr);

#ifdef __cplusplus
}
#endif
#endif
