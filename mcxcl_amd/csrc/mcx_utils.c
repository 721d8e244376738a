/*
 * mcx_utils.c - configuration layer: defaults, JSON and legacy .inp readers,
 * volume loading, input validation and preprocessing, detector masking,
 * replay preparation, normalisation helpers and output writers.
 *
 * Behaviour follows the reference configuration layer (src/mcx_utils.c); each
 * function cites the reference function it mirrors.  The JSON reader is the
 * project's own (mcx_json.c) instead of the vendored cJSON.
 */
#define _GNU_SOURCE
#include <ctype.h>
#include <math.h>
#include <setjmp.h>
#include <stddef.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "mcx_jdata.h"
#include "mcx_json.h"
#include "mcx_shapes.h"
#include "mcx_utils.h"

#define MIN(a, b) ((a) < (b) ? (a) : (b))
#define MAX(a, b) ((a) > (b) ? (a) : (b))

/* ---------------- name tables (reference src/mcx_utils.c:132-199) ---------------- */
const char* srctypeid[] = {"pencil", "isotropic", "cone", "gaussian", "planar", "pattern", "fourier", "arcsine",
                           "disk", "fourierx", "fourierx2d", "zgaussian", "line", "slit", "pencilarray",
                           "pattern3d", "hyperboloid", "ring", ""
                          };
const char outputtype[] = {'x', 'f', 'e', 'j', 'p', 'm', 'r', 'l', 's', 't', 'b', 'a', 'd', 'u', 'v', 'w', 'q', '\0'};
const char boundarycond[] = {'_', 'r', 'a', 'm', 'c', '\0'};
const char boundarydetflag[] = {'0', '1', '\0'};
const char saveflag[] = {'D', 'S', 'P', 'M', 'X', 'V', 'W', 'I', '\0'};
const char debugflag[] = {'R', 'M', 'P', 'T', '\0'};
static const char* outputformat_names[] = {"mc2", "nii", "hdr", "ubj", "tx3", "jnii", "bnii", ""};
static const unsigned int mediaformatid[] = {1, 2, 4, 96, 97, 98, 99, 100, 101, 102, 103, 104,
                                             MEDIA_MUA_MUS_FLOAT, 0
                                            };
static const char* mediaformat_names[] = {"byte", "short", "integer", "asgn_float", "svmc", "mixlabel", "labelplus",
                                          "muamus_float", "mua_float", "muamus_half", "asgn_byte", "muamus_short",
                                          "muamus_f32", ""
                                         };

/* command-line flags that take precedence over JSON values (reference flagset[]) */
static char flagset[256];

/* ---------------- error handling ---------------- */
static __thread jmp_buf* mcx_errjmp = NULL;
static __thread char mcx_errmsg[2048];

const char* mcx_last_error(void) {
    return mcx_errmsg;
}

void* mcx_push_errjmp(void* jb) {
    void* old = (void*)mcx_errjmp;
    mcx_errjmp = (jmp_buf*)jb;
    return old;
}

void mcx_pop_errjmp(void* old) {
    mcx_errjmp = (jmp_buf*)old;
}

/* reference src/mcx_utils.c:1298-1305 */
void mcx_error(const int id, const char* msg, const char* file, const int linenum) {
    snprintf(mcx_errmsg, sizeof(mcx_errmsg), "MCX ERROR(%d):%s in unit %s:%d", id, msg, file, linenum);

    if (mcx_errjmp) {
        longjmp(*mcx_errjmp, id ? id : -1);
    }

    fprintf(stdout, "\n%s\n", mcx_errmsg);
    fflush(stdout);
    exit(id);
}

void mcx_assess(const int err, const char* msg, const char* file, const int linenum) {
    if (!err) {
        mcx_error(err, msg, file, linenum);
    }
}

void mcx_flush(Config* cfg) {
    fflush(cfg->flog ? cfg->flog : stdout);
}

/* ---------------- defaults (reference src/mcx_utils.c:203-366) ---------------- */
void mcx_initcfg(Config* cfg) {
    memset(cfg, 0, sizeof(Config));
    memset(flagset, 0, sizeof(flagset));   /* command-line overrides belong to one configuration */
    cfg->tstart = 0.f;
    cfg->tend = 5e-9f;
    cfg->tstep = 5e-9f;
    cfg->mediabyte = 1;
    cfg->steps.x = cfg->steps.y = cfg->steps.z = 1.f;
    cfg->nblocksize = 64;
    cfg->nthread = (1 << 14);
    cfg->seed = 0x623F9A9E;
    cfg->isreflect = 1;
    cfg->isnormalized = 1;
    cfg->issavedet = 1;
    cfg->respin = 1;
    cfg->issave2pt = 1;
    cfg->unitinmm = 1.f;
    cfg->flog = stdout;
    cfg->savedetflag = 0x5;
    cfg->his.savedetflag = cfg->savedetflag;
    cfg->gscatter = 1000000000u;
    cfg->maxvoidstep = 1000;
    cfg->voidtime = 1;
    cfg->maxdetphoton = 1000000;
    cfg->autopilot = 1;
    cfg->optlevel = 3;
    cfg->srcnum = 1;
    cfg->deviceid[0] = '1';
    cfg->maxjumpdebug = 10000000;
    cfg->isatomic = 1;
    cfg->srcpos.w = 1.f;
    memcpy(cfg->his.magic, "MCXH", 4);
    cfg->his.version = 1;
    cfg->his.unitinmm = 1.f;
    cfg->his.normalizer = 1.f;
    cfg->his.srcnum = 1;
    cfg->his.respin = 1;
    cfg->parentid = mpStandalone;
    cfg->outputtype = otFlux;
    cfg->outputformat = ofJNifti;
    cfg->srciquv.x = 1.f;
    cfg->his.totalsource = 1;
    cfg->mathmode = 0;
}

/* reference src/mcx_utils.c:368-469 */
void mcx_clearcfg(Config* cfg) {
    void* ptrs[] = {cfg->prop, cfg->polprop, cfg->smatrix, cfg->detpos, cfg->detdir, cfg->vol, cfg->srcpattern,
                    cfg->replay.weight, cfg->replay.seed, cfg->replay.tof, cfg->replay.detid, cfg->exportfield,
                    cfg->exportjacob, cfg->exportdetected, cfg->exportdebugdata, cfg->seeddata, cfg->shapedata,
                    cfg->extrajson, cfg->invcdf, cfg->angleinvcdf, cfg->srcdata
                   };

    for (size_t i = 0; i < sizeof(ptrs) / sizeof(ptrs[0]); i++) {
        free(ptrs[i]);
    }

    FILE* flog = cfg->flog;
    mcx_initcfg(cfg);

    if (flog && flog != stdout && flog != stderr) {
        fclose(flog);
    }
}

/* ---------------- small lookups (reference src/mcx_utils.c:5207-5305) ---------------- */
int mcx_lookupindex(char* key, const char* index) {
    for (int i = 0; index[i] != '\0'; i++) {
        if (tolower((unsigned char)*key) == index[i]) {
            *key = (char)i;
            return 0;
        }
    }

    return 1;
}

int mcx_keylookup(char* origkey, const char* table[]) {
    char key[256];
    size_t n = strlen(origkey);

    if (n >= sizeof(key)) {
        return -1;
    }

    for (size_t i = 0; i <= n; i++) {
        key[i] = (char)tolower((unsigned char)origkey[i]);
    }

    for (int i = 0; table[i] && table[i][0]; i++) {
        if (!strcmp(key, table[i])) {
            return i;
        }
    }

    return -1;
}

int mcx_parsedebugopt(char* debugopt, const char* flags) {
    int level = 0;

    for (const char* c = debugopt; *c; c++) {
        char u = (*c >= 'a' && *c <= 'z') ? (char)(*c - 'a' + 'A') : *c;
        const char* p = strchr(flags, u);

        if (p && u) {
            level |= (1 << (p - flags));
        }
    }

    return level;
}

/* reference src/mcx_host.cpp:494-496 */
unsigned int mcx_partialdata(const Config* cfg) {
    return (cfg->medianum - 1) * (SAVE_NSCAT(cfg->savedetflag) + SAVE_PPATH(cfg->savedetflag) +
                                  SAVE_MOM(cfg->savedetflag));
}

unsigned int mcx_detreclen(const Config* cfg) {
    return mcx_partialdata(cfg) + SAVE_DETID(cfg->savedetflag) +
           3 * (SAVE_PEXIT(cfg->savedetflag) + SAVE_VEXIT(cfg->savedetflag)) + SAVE_W0(cfg->savedetflag) +
           4 * SAVE_IQUV(cfg->savedetflag);
}

/* ---------------- numeric helpers (reference src/mcx_utils.c:1208-1273) ---------------- */
void mcx_normalize(float field[], float scale, int fieldlen, int option, int pidx, int srcnum) {
    for (int i = 0; i < fieldlen; i++) {
        if (option == 2 && field[i * srcnum + pidx] < 0.f) {
            continue;
        }

        field[i * srcnum + pidx] *= scale;
    }
}

void mcx_kahanSum(float* sum, float* kahanc, float input) {
    float y = input - *kahanc;
    float t = *sum + y;
    *kahanc = (t - *sum) - y;
    *sum = t;
}

float mcx_updatemua(unsigned int mediaid, Config* cfg) {
    float mua = 0.f;

    if (cfg->mediabyte <= 4) {
        mua = cfg->prop[mediaid & MED_MASK].mua;
    } else if (cfg->mediabyte == MEDIA_MUA_FLOAT) {
        union {
            float f;
            unsigned int i;
        } m;
        m.i = mediaid;
        mua = fabsf(m.f);
    } else if (cfg->mediabyte == MEDIA_ASGN_BYTE) {
        mua = (mediaid & 0xFF) * (1.f / 255.f) * (cfg->prop[2].mua - cfg->prop[1].mua) + cfg->prop[1].mua;
    } else if (cfg->mediabyte == MEDIA_AS_SHORT) {
        mua = (mediaid & 0xFFFF) * (1.f / 65535.f) * (cfg->prop[2].mua - cfg->prop[1].mua) + cfg->prop[1].mua;
    }

    return mua;
}

void mcx_createfluence(float** fluence, Config* cfg) {
    mcx_clearfluence(fluence);
    *fluence = (float*)calloc((size_t)cfg->dim.x * cfg->dim.y * cfg->dim.z, cfg->maxgate * sizeof(float));
}

void mcx_clearfluence(float** fluence) {
    if (*fluence) {
        free(*fluence);
        *fluence = NULL;
    }
}

void mcx_cleargpuinfo(GPUInfo** gpuinfo) {
    if (*gpuinfo) {
        free(*gpuinfo);
        *gpuinfo = NULL;
    }
}

/* reference src/mcx_utils.c:3969-4083: C (row-major) to Fortran (column-major) */
void mcx_convertrow2col(unsigned int** vol, MCXUInt4* dim) {
    size_t nx = dim->x, ny = dim->y, nz = dim->z;
    unsigned int* out = (unsigned int*)malloc(nx * ny * nz * sizeof(unsigned int));

    for (size_t i = 0; i < nx; i++)
        for (size_t j = 0; j < ny; j++)
            for (size_t k = 0; k < nz; k++) {
                out[k * nx * ny + j * nx + i] = (*vol)[i * ny * nz + j * nz + k];
            }

    free(*vol);
    *vol = out;
}

/* the same for volumes of 8 bytes per voxel (svmc and muamus_float as read,
   reference src/mcx_utils.c:1606-1608) */
static void mcx_convertrow2col64(unsigned int** vol, MCXUInt4* dim) {
    size_t nx = dim->x, ny = dim->y, nz = dim->z;
    unsigned long long* in = (unsigned long long*)(*vol);
    unsigned long long* out = (unsigned long long*)malloc(nx * ny * nz * sizeof(unsigned long long));

    for (size_t i = 0; i < nx; i++)
        for (size_t j = 0; j < ny; j++)
            for (size_t k = 0; k < nz; k++) {
                out[k * nx * ny + j * nx + i] = in[i * ny * nz + j * nz + k];
            }

    free(*vol);
    *vol = (unsigned int*)out;
}

/* split-voxel volumes (svmc): 8 bytes per voxel as read, {lower label, upper
   label, reference point x, y, z, normal x, y, z}, become two blocks of nvox
   words: the first holds py | px << 8 | upper << 16 | lower << 24 (bit 31 is
   also the detector flag), the second nz | ny << 8 | nx << 16 | pz << 24
   (reference src/mcx_utils.c:1688-1712) */
static void mcx_split_svmc(Config* cfg) {
    const size_t n = (size_t)cfg->dim.x * cfg->dim.y * cfg->dim.z;
    const unsigned char* val = (const unsigned char*)cfg->vol;
    unsigned int* out = (unsigned int*)malloc(n * 2 * sizeof(unsigned int));

    for (size_t i = 0; i < n; i++) {
        const unsigned char* b = val + (i << 3);
        out[i] = (unsigned int)b[3] | ((unsigned int)b[2] << 8) | ((unsigned int)b[1] << 16) |
                 ((unsigned int)b[0] << 24);
        out[i + n] = (unsigned int)b[7] | ((unsigned int)b[6] << 8) | ((unsigned int)b[5] << 16) |
                     ((unsigned int)b[4] << 24);
    }

    free(cfg->vol);
    cfg->vol = out;
}

/* ---------------- packed media words ---------------- */

/* float32 bits to the binary16 pattern the reference's packers write: the
   exponent rebiased and clamped at zero (wrapping above the half range), the
   mantissa truncated.  With withsub, a value whose rebiased exponent is zero
   and whose top ten mantissa bits are below 0x10 is rounded to a subnormal
   half (src/mcx_utils.c:3625-3646); the labelplus packer has no such path
   (src/pmcxcl.cpp:283-288).  A right shift of 32 or more counts as zero: such
   values are below 2^-45 and the reference's shift there is undefined. */
static unsigned short mcx_half_bits(unsigned int i, int withsub) {
    const unsigned int m = (i >> 13) & 0x3FFu, e = (i >> 23) & 0xFFu;
    const unsigned int t = (e > 0x70u) ? ((e - 0x70u) & 0x1Fu) : 0u;

    if (withsub && m < 0x10u && t == 0) {
        const unsigned int mm = ((i >> 12) & 0x7FFu) | 0x800u;
        const unsigned int s1 = 114u - e, s0 = 113u - e;   /* e <= 0x70 here: both positive */
        const unsigned int a = (s1 < 32u) ? (mm >> s1) : 0u;
        const unsigned int b = (s0 < 32u) ? ((mm >> s0) & 1u) : 0u;
        return (unsigned short)(((i >> 16) & 0x8000u) | ((a + b) & 0xFFFFu));
    }

    return (unsigned short)(((i >> 31) << 15) | (t << 10) | m);
}

/* Pack per-voxel properties into the words of a packed media format, as the
   reference's volume reader (src/mcx_utils.c:3597-3681) and its Python module
   (src/pmcxcl.cpp:243-389) do.  in holds ncomp floats per voxel, voxel-major:
     MEDIA_MUA_FLOAT  (1): mua (1/mm) x unitinmm as float32 bits; a zero word
                           becomes EPS, NaN and -0 become 0 (void)
     MEDIA_AS_F2H     (2): {mua, mus} (1/mm) x unitinmm as two halves, mua in the
                           low half; NaN in either gives 0, an all-zero word
                           0x00010000 (mus = the smallest half), -0 gives 0
     MEDIA_LABEL_HALF (3): {value, component 0..3, label}: label | component << 14
                           in the low half, the value's half in the high half;
                           the value is scaled by unitinmm for components 1 and 2,
                           as src/pmcxcl.cpp:273-275 does
   Returns 0; -1 for another format; -2 for a labelplus component outside 0..3
   or a label outside 0..16383 (the reference truncates such labels into the
   component bits). */
int mcx_pack_media(unsigned int format, const float* in, size_t nvox, float unitinmm, unsigned int* out) {
    union {
        float f;
        unsigned int u;
    } a, b;

    if (format == MEDIA_MUA_FLOAT) {
        for (size_t i = 0; i < nvox; i++) {
            a.f = in[i] * unitinmm;

            if (a.u == 0) {
                a.f = EPS;
            }

            if (in[i] != in[i] || a.u == SIGN_BIT) {
                a.u = 0;
            }

            out[i] = a.u;
        }
    } else if (format == MEDIA_AS_F2H) {
        for (size_t i = 0; i < nvox; i++) {
            a.f = in[2 * i] * unitinmm;
            b.f = in[2 * i + 1] * unitinmm;

            if (a.f != a.f || b.f != b.f) {
                out[i] = 0;
                continue;
            }

            unsigned int w = (unsigned int)mcx_half_bits(a.u, 1) | ((unsigned int)mcx_half_bits(b.u, 1) << 16);
            w = (w == 0) ? 0x00010000u : w;
            out[i] = (w == SIGN_BIT) ? 0u : w;
        }
    } else if (format == MEDIA_LABEL_HALF) {
        for (size_t i = 0; i < nvox; i++) {
            a.f = in[3 * i];
            const float comp = in[3 * i + 1], label = in[3 * i + 2];

            if (!(comp >= 0.f && comp < 4.f) || !(label >= 0.f && label < 16384.f)) {
                return -2;
            }

            if (comp >= 1.f && comp <= 2.f) {
                a.f *= unitinmm;
            }

            out[i] = ((((unsigned int)comp & 3u) << 14) | (unsigned int)label) |
                     ((unsigned int)mcx_half_bits(a.u, 0) << 16);
        }
    } else {
        return -1;
    }

    return 0;
}

/* ---------------- volume file (reference src/mcx_utils.c:3524-3698) ---------------- */
static void mcx_loadvolume(const char* filename, Config* cfg) {
    size_t n = (size_t)cfg->dim.x * cfg->dim.y * cfg->dim.z;
    FILE* fp = fopen(filename, "rb");

    if (!fp) {
        MCX_ERROR(-5, "the specified binary volume file does not exist");
    }

    unsigned int bytes = (cfg->mediabyte <= 4) ? cfg->mediabyte : 4;
    /* two words per voxel kept as read */
    int pairs = (cfg->mediabyte == MEDIA_2LABEL_SPLIT || cfg->mediabyte == MEDIA_MUA_MUS_FLOAT);

    if (pairs || cfg->mediabyte == MEDIA_AS_F2H) {
        bytes = 8;
    }

    unsigned char* raw = (unsigned char*)malloc(n * bytes);
    size_t got = fread(raw, bytes, n, fp);
    fclose(fp);

    if (got != n) {
        free(raw);
        MCX_ERROR(-6, "file size does not match specified dimensions");
    }

    free(cfg->vol);
    cfg->vol = (unsigned int*)malloc(n * sizeof(unsigned int) * (pairs ? 2 : 1));

    if (cfg->mediabyte == 1) {
        for (size_t i = 0; i < n; i++) {
            cfg->vol[i] = raw[i];
        }
    } else if (cfg->mediabyte == 2) {
        for (size_t i = 0; i < n; i++) {
            cfg->vol[i] = ((unsigned short*)raw)[i];
        }
    } else if (cfg->mediabyte == MEDIA_MUA_FLOAT || cfg->mediabyte == MEDIA_AS_F2H) {
        /* float mua, or {mua, mus} pairs, in 1/mm: packed into one word per voxel */
        mcx_pack_media(cfg->mediabyte, (const float*)raw, n, cfg->unitinmm, cfg->vol);
    } else {
        memcpy(cfg->vol, raw, n * bytes);
    }

    free(raw);
}

/* ---------------- JSON reader (reference src/mcx_utils.c:2401-3286) ---------------- */
/* look a key up in the section, falling back to a flattened "Section.Key" at the root */
static MJNode* jfind(MJNode* root, MJNode* parent, const char* id, const char* idfull) {
    MJNode* v = mj_get(parent, id);

    if (!v && idfull) {
        v = mj_get(root, idfull);
    }

    return v;
}

static void read_quad(MJNode* arr, MCXFloat4* out, int minlen) {
    float* f = &out->x;

    for (int i = 0; i < 4 && i < arr->count; i++) {
        MJNode* v = mj_at(arr, i);

        if (v->type == MJ_STRING) {
            if (!strcmp(v->str, "_NaN_")) {
                f[i] = NAN;
            } else if (!strcmp(v->str, "_Inf_")) {
                f[i] = INFINITY;
            } else if (!strcmp(v->str, "-_Inf_")) {
                f[i] = -INFINITY;
            }
        } else {
            f[i] = (float)mj_num(v, 0.0);
        }
    }

    (void)minlen;
}

/* Source Pos/Dir/Param1/Param2 may be a single vector or a list of vectors
   (first one is the primary source, the rest are extra sources) */
static void read_src_vector(Config* cfg, MJNode* item, int which) {
    if (!item || item->type != MJ_ARRAY || item->count == 0) {
        return;
    }

    MJNode* first = item;

    if (item->child->type == MJ_ARRAY) {
        if (cfg->srcdata && cfg->extrasrclen != (unsigned int)(item->count - 1)) {
            MCX_ERROR(-1, "Length of sub-elements of Pos/Dir/Param1/Param2 must match");
        }

        if (!cfg->srcdata) {
            cfg->extrasrclen = item->count - 1;
        }

        first = item->child;
    }

    MCXFloat4* dst = (which == 0) ? &cfg->srcpos : (which == 1) ? &cfg->srcdir : (which == 2) ? &cfg->srcparam1 :
                     &cfg->srcparam2;

    if (which == 0) {
        MCXFloat4 v = *dst;
        read_quad(first, &v, 3);
        dst->x = v.x;
        dst->y = v.y;
        dst->z = v.z;

        if (first->count > 3) {
            dst->w = v.w;
        }
    } else if (which == 1) {
        MCXFloat4 v = *dst;
        read_quad(first, &v, 3);
        *dst = v;
    } else {
        read_quad(first, dst, 1);
    }

    if (cfg->extrasrclen > 0 && first != item) {
        if (!cfg->srcdata) {
            cfg->srcdata = (ExtraSrc*)calloc(cfg->extrasrclen, sizeof(ExtraSrc));
        }

        unsigned int count = 0;

        for (MJNode* s = first->next; s && count < cfg->extrasrclen; s = s->next, count++) {
            ExtraSrc* es = cfg->srcdata + count;

            if (which == 0) {
                MCXFloat4 v = es->srcpos;
                read_quad(s, &v, 3);
                es->srcpos.x = v.x - (!cfg->issrcfrom0);
                es->srcpos.y = v.y - (!cfg->issrcfrom0);
                es->srcpos.z = v.z - (!cfg->issrcfrom0);

                if (s->count > 3) {
                    es->srcpos.w = v.w;
                }
            } else if (which == 1) {
                read_quad(s, &es->srcdir, 3);
            } else if (which == 2) {
                read_quad(s, &es->srcparam1, 1);
            } else {
                read_quad(s, &es->srcparam2, 1);
            }
        }
    }
}

static int mcx_loadjson(MJNode* root, Config* cfg) {
    MJNode* Domain = mj_get(root, "Domain");
    MJNode* Optode = mj_get(root, "Optode");
    MJNode* Session = mj_get(root, "Session");
    MJNode* Forward = mj_get(root, "Forward");
    MJNode* Shapes = mj_get(root, "Shapes");
    char filename[MAX_FULL_PATH] = {'\0'};
    MJNode* v;

    if (Domain) {
        v = jfind(root, Domain, "VolumeFile", "Domain.VolumeFile");

        if (v && v->type == MJ_STRING) {
            if (cfg->rootpath[0]) {
                snprintf(filename, sizeof(filename), "%s/%s", cfg->rootpath, v->str);
            } else {
                snprintf(filename, sizeof(filename), "%s", v->str);
            }
        }

        if (!flagset['K']) {
            v = jfind(root, Domain, "MediaFormat", "Domain.MediaFormat");

            if (v) {
                int id = mcx_keylookup((char*)mj_str(v, "byte"), mediaformat_names);

                if (id < 0) {
                    MCX_ERROR(-1, "Unsupported media format");
                }

                cfg->mediabyte = mediaformatid[id];
            }
        }

        if (!flagset['u']) {
            cfg->unitinmm = (float)mj_num(jfind(root, Domain, "LengthUnit", "Domain.LengthUnit"), 1.0);
        }

        MJNode* meds = jfind(root, Domain, "Media", "Domain.Media");

        if (meds && meds->count > 0) {
            cfg->medianum = meds->count;

            if (cfg->medianum > MAX_PROP) {
                MCX_ERROR(-4, "input media types exceed the maximum");
            }

            free(cfg->prop);
            cfg->prop = (Medium*)calloc(cfg->medianum, sizeof(Medium));
            unsigned int i = 0;

            for (MJNode* m = meds->child; m && i < cfg->medianum; m = m->next, i++) {
                if (m->type == MJ_OBJECT) {
                    const char* keys[4] = {"mua", "mus", "g", "n"};
                    float* dst = &cfg->prop[i].mua;

                    for (int k = 0; k < 4; k++) {
                        MJNode* x = mj_get(m, keys[k]);

                        if (!x) {
                            MCX_ERROR(-1, "Each medium needs mua, mus, g and n");
                        }

                        dst[k] = (float)mj_num(x, 0.0);
                    }
                } else if (m->type == MJ_ARRAY && m->count >= 4) {
                    cfg->prop[i].mua = (float)mj_num(mj_at(m, 0), 0);
                    cfg->prop[i].mus = (float)mj_num(mj_at(m, 1), 0);
                    cfg->prop[i].g = (float)mj_num(mj_at(m, 2), 0);
                    cfg->prop[i].n = (float)mj_num(mj_at(m, 3), 0);
                } else {
                    MCX_ERROR(-1, "Session.Media must be either an array of objects or array of 4-elem numerical arrays");
                }
            }
        }

        /* polarised media (src/mcx_utils.c:2510-2566): {mua, radius, rho, nsph, nmed}
           objects or 5-element arrays, one per medium from 1 */
        MJNode* mies = jfind(root, Domain, "MieScatter", "Domain.MieScatter");

        if (mies && mies->count > 0) {
            const char* keys[5] = {"mua", "radius", "rho", "nsph", "nmed"};
            unsigned int i = 0;
            cfg->polmedianum = mies->count;
            free(cfg->polprop);
            cfg->polprop = (POLMedium*)calloc(cfg->polmedianum, sizeof(POLMedium));

            for (MJNode* m = mies->child; m && i < cfg->polmedianum; m = m->next, i++) {
                float* dst = &cfg->polprop[i].mua;

                for (int k = 0; k < 5; k++) {
                    MJNode* x = (m->type == MJ_OBJECT) ? mj_get(m, keys[k]) :
                                ((m->type == MJ_ARRAY && m->count >= 5) ? mj_at(m, k) : NULL);

                    if (x) {
                        dst[k] = (float)mj_num(x, 0.0);
                    }
                }
            }
        }

        v = jfind(root, Domain, "Dim", "Domain.Dim");

        if (v && v->count >= 3) {
            cfg->dim.x = (unsigned int)mj_int(mj_at(v, 0), 0);
            cfg->dim.y = (unsigned int)mj_int(mj_at(v, 1), 0);
            cfg->dim.z = (unsigned int)mj_int(mj_at(v, 2), 0);
        } else if (!Shapes) {
            MCX_ERROR(-1, "You must specify the dimension of the volume");
        }

        v = jfind(root, Domain, "Step", "Domain.Step");

        if (v) {
            if (v->count < 3) {
                MCX_ERROR(-1, "Domain::Step has incorrect element numbers");
            }

            cfg->steps.x = (float)mj_num(mj_at(v, 0), 1);
            cfg->steps.y = (float)mj_num(mj_at(v, 1), 1);
            cfg->steps.z = (float)mj_num(mj_at(v, 2), 1);
        }

        if (cfg->steps.x != cfg->steps.y || cfg->steps.y != cfg->steps.z) {
            MCX_ERROR(-9, "MCX currently does not support anisotropic voxels");
        }

        if (cfg->steps.x != 1.f && cfg->unitinmm == 1.f) {
            cfg->unitinmm = cfg->steps.x;
        }

        if (cfg->unitinmm != 1.f) {
            cfg->steps.x = cfg->steps.y = cfg->steps.z = cfg->unitinmm;
        }

        v = jfind(root, Domain, "InverseCDF", "Domain.InverseCDF");

        if (v && v->type == MJ_ARRAY) {
            int np = v->count;
            cfg->nphase = np + 2;
            free(cfg->invcdf);
            cfg->invcdf = (float*)calloc(cfg->nphase, sizeof(float));
            cfg->invcdf[0] = -1.f;
            int i = 1;

            for (MJNode* x = v->child; x; x = x->next, i++) {
                cfg->invcdf[i] = (float)mj_num(x, 0);

                if (cfg->invcdf[i] < cfg->invcdf[i - 1] || cfg->invcdf[i] > 1.f || cfg->invcdf[i] < -1.f) {
                    MCX_ERROR(-1, "Domain.InverseCDF contains invalid data; it must be a monotonically increasing vector with all values between -1 and 1");
                }
            }

            cfg->invcdf[cfg->nphase - 1] = 1.f;
        }

        v = jfind(root, Domain, "OriginType", "Domain.OriginType");

        if (v && cfg->issrcfrom0 == 0) {
            cfg->issrcfrom0 = (char)mj_int(v, 0);
        }
    }

    if (Optode) {
        MJNode* src = jfind(root, Optode, "Source", "Optode.Source");

        if (src) {
            v = jfind(root, src, "Pos", "Optode.Source.Pos");

            if (v && v->type == MJ_ARRAY) {
                cfg->extrasrclen = 0;
                read_src_vector(cfg, v, 0);
            }

            v = jfind(root, src, "Weight", "Optode.Source.Weight");

            if (v) {
                cfg->srcpos.w = (float)mj_num(v, cfg->srcpos.w);
            }

            read_src_vector(cfg, jfind(root, src, "Dir", "Optode.Source.Dir"), 1);

            if (!cfg->issrcfrom0) {
                cfg->srcpos.x--;
                cfg->srcpos.y--;
                cfg->srcpos.z--;
            }

            v = jfind(root, src, "Type", "Optode.Source.Type");

            if (v) {
                cfg->srctype = (char)mcx_keylookup((char*)mj_str(v, "pencil"), srctypeid);
            }

            read_src_vector(cfg, jfind(root, src, "Param1", "Optode.Source.Param1"), 2);
            read_src_vector(cfg, jfind(root, src, "Param2", "Optode.Source.Param2"), 3);

            v = jfind(root, src, "Frequency", "Optode.Source.Frequency");

            if (v) {
                cfg->omega = (float)(mj_num(v, 0) * TWO_PI);
            }

            v = jfind(root, src, "WaveLength", "Optode.Source.WaveLength");

            if (v) {
                cfg->lambda = (float)mj_num(v, 0);
            }

            v = jfind(root, src, "SrcNum", "Optode.Source.SrcNum");

            if (v) {
                cfg->srcnum = (unsigned int)mj_int(v, (int)cfg->srcnum);
            }

            v = jfind(root, src, "ID", "Optode.Source.ID");

            if (v) {
                cfg->srcid = mj_int(v, cfg->srcid);
            }

            MJNode* pat = jfind(root, src, "Pattern", "Optode.Source.Pattern");

            if (pat) {
                int nx = mj_int(mj_get(pat, "Nx"), 0), ny = mj_int(mj_get(pat, "Ny"), 0);
                int nz = mj_int(mj_get(pat, "Nz"), 1);
                MJNode* data = mj_get(pat, "Data");

                if (nx > 0 && ny > 0 && data) {
                    size_t len = (size_t)nx * ny * nz * cfg->srcnum;
                    free(cfg->srcpattern);
                    cfg->srcpattern = (float*)calloc(len, sizeof(float));

                    if (data->type == MJ_ARRAY) {
                        MJNode* x = data->child;

                        for (size_t i = 0; i < len; i++, x = x ? x->next : NULL) {
                            if (!x) {
                                MCX_ERROR(-1, "Incomplete pattern data");
                            }

                            cfg->srcpattern[i] = (float)mj_num(x, 0);
                        }
                    } else if (data->type == MJ_STRING) {
                        FILE* fp = fopen(data->str, "rb");

                        if (fp) {
                            if (fread(cfg->srcpattern, sizeof(float), len, fp) != len) {
                                fclose(fp);
                                MCX_ERROR(-2, "srcpattern file size does not match the specified data size");
                            }

                            fclose(fp);
                        }
                    }
                }
            }

            v = jfind(root, src, "AngleInverseCDF", "Optode.Source.AngleInverseCDF");

            if (v && v->type == MJ_ARRAY) {
                cfg->nangle = v->count;
                free(cfg->angleinvcdf);
                cfg->angleinvcdf = (float*)calloc(cfg->nangle, sizeof(float));
                int i = 0;

                for (MJNode* x = v->child; x; x = x->next, i++) {
                    cfg->angleinvcdf[i] = (float)mj_num(x, 0);

                    if ((i > 0 && cfg->angleinvcdf[i] < cfg->angleinvcdf[i - 1]) || cfg->angleinvcdf[i] > 1.f ||
                            cfg->angleinvcdf[i] < 0.f) {
                        MCX_ERROR(-1, "Optode.Source.AngleInverseCDF contains invalid data; it must be a monotonically increasing vector with all values between 0 and 1");
                    }
                }
            }
        }

        MJNode* dets = jfind(root, Optode, "Detector", "Optode.Detector");

        if (dets && dets->count > 0) {
            cfg->detnum = dets->count;
            free(cfg->detpos);
            cfg->detpos = (MCXFloat4*)calloc(cfg->detnum, sizeof(MCXFloat4));
            unsigned int i = 0;

            for (MJNode* d = dets->child; d && i < cfg->detnum; d = d->next, i++) {
                MJNode* pos = mj_get(d, "Pos");
                MJNode* rad = mj_get(d, "R");
                MJNode* dir = mj_get(d, "Dir");

                if (pos && pos->count >= 3) {
                    cfg->detpos[i].x = (float)mj_num(mj_at(pos, 0), 0);
                    cfg->detpos[i].y = (float)mj_num(mj_at(pos, 1), 0);
                    cfg->detpos[i].z = (float)mj_num(mj_at(pos, 2), 0);
                }

                if (rad) {
                    cfg->detpos[i].w = (float)mj_num(rad, 0);
                }

                if (dir && dir->type == MJ_ARRAY && dir->count >= 3) {
                    if (!cfg->detdir) {
                        cfg->detdir = (MCXFloat4*)calloc(cfg->detnum, sizeof(MCXFloat4));
                    }

                    read_quad(dir, cfg->detdir + i, 3);
                }

                if (!cfg->issrcfrom0) {
                    cfg->detpos[i].x--;
                    cfg->detpos[i].y--;
                    cfg->detpos[i].z--;
                }
            }
        }
    }

    if (cfg->medianum + cfg->detnum + (cfg->extrasrclen << 2) > MAX_PROP) {
        MCX_ERROR(-4, "input media types plus detector number plus additional sources exceeds the maximum total");
    }

    if (Session) {
        if (!flagset['E']) {
            cfg->seed = mj_int(jfind(root, Session, "RNGSeed", "Session.RNGSeed"), -1);
        }

        if (!flagset['n']) {
            cfg->nphoton = (size_t)mj_num(jfind(root, Session, "Photons", "Session.Photons"), 0);
        }

        if (!flagset['H']) {
            cfg->maxdetphoton = (unsigned int)mj_num(jfind(root, Session, "MaxDetPhoton", "Session.MaxDetPhoton"),
                                cfg->maxdetphoton);
        }

        if (cfg->session[0] == '\0') {
            snprintf(cfg->session, MAX_SESSION_LENGTH, "%s", mj_str(jfind(root, Session, "ID", "Session.ID"), "default"));
        }

        if (cfg->rootpath[0] == '\0') {
            snprintf(cfg->rootpath, MAX_PATH_LENGTH, "%s", mj_str(jfind(root, Session, "RootPath", "Session.RootPath"), ""));
        }

        if (!flagset['B']) {
            const char* bc = mj_str(jfind(root, Session, "BCFlags", "Session.BCFlags"), NULL);

            if (bc) {
                memset(cfg->bc, 0, sizeof(cfg->bc));
                strncpy(cfg->bc, bc, 12);
            }
        }

#define SESSION_FLAG(letter, key, field) \
        if (!flagset[letter]) { v = jfind(root, Session, key, "Session." key); if (v) cfg->field = (char)mj_int(v, cfg->field); }
        SESSION_FLAG('b', "DoMismatch", isreflect);
        SESSION_FLAG('S', "DoSaveVolume", issave2pt);
        SESSION_FLAG('U', "DoNormalize", isnormalized);
        SESSION_FLAG('d', "DoPartialPath", issavedet);
        SESSION_FLAG('X', "DoSaveRef", issaveref);
        SESSION_FLAG('x', "DoSaveExit", issaveexit);
        SESSION_FLAG('q', "DoSaveSeed", issaveseed);
        SESSION_FLAG('A', "DoAutoThread", autopilot);
        SESSION_FLAG('m', "DoDCS", ismomentum);
        SESSION_FLAG('V', "DoSpecular", isspecular);
#undef SESSION_FLAG

        if (!flagset['D']) {
            v = jfind(root, Session, "DebugFlag", "Session.DebugFlag");

            if (v && v->type == MJ_STRING) {
                cfg->debuglevel = (unsigned int)mcx_parsedebugopt(v->str, debugflag);
            } else if (v) {
                cfg->debuglevel = (unsigned int)mj_int(v, 0);
            }
        }

        if (!flagset['w']) {
            v = jfind(root, Session, "SaveDataMask", "Session.SaveDataMask");

            if (v && v->type == MJ_STRING) {
                cfg->savedetflag = (unsigned int)mcx_parsedebugopt(v->str, saveflag);
            } else if (v) {
                cfg->savedetflag = (unsigned int)mj_int(v, 5);
            }
        }

        if (!flagset['F']) {
            v = jfind(root, Session, "OutputFormat", "Session.OutputFormat");

            if (v) {
                int id = mcx_keylookup((char*)mj_str(v, "jnii"), outputformat_names);

                if (id < 0) {
                    MCX_ERROR(-2, "the specified output format is not recognized");
                }

                cfg->outputformat = (char)id;
            }
        }

        v = jfind(root, Session, "OutputType", "Session.OutputType");

        if (v && v->type == MJ_STRING && v->str[0]) {
            char c = v->str[0];

            if (mcx_lookupindex(&c, outputtype)) {
                MCX_ERROR(-2, "the specified output data type is not recognized");
            }

            if (!flagset['O']) {
                cfg->outputtype = c;
            }
        }

        if (!flagset['e']) {
            cfg->minenergy = (float)mj_num(jfind(root, Session, "MinEnergy", "Session.MinEnergy"), cfg->minenergy);
        }
    }

    if (Forward) {
        cfg->tstart = (float)mj_num(jfind(root, Forward, "T0", "Forward.T0"), 0.0);
        cfg->tend = (float)mj_num(jfind(root, Forward, "T1", "Forward.T1"), 0.0);
        cfg->tstep = (float)mj_num(jfind(root, Forward, "Dt", "Forward.Dt"), 0.0);

        if (cfg->tstart > cfg->tend || cfg->tstep == 0.f) {
            MCX_ERROR(-9, "incorrect time gate settings");
        }

        unsigned int gates = (unsigned int)((cfg->tend - cfg->tstart) / cfg->tstep + 0.5);

        if (cfg->maxgate == 0 || cfg->maxgate > gates) {
            cfg->maxgate = gates;
        }
    }

    if (filename[0] == '\0') {
        if (Shapes && mj_get(Shapes, "_ArraySize_")) {
            /* a label volume embedded as a JData array (src/mcx_utils.c:3239-3250) */
            unsigned int dims[4] = {1, 1, 1, 1};
            int ndim = 0;
            char type[32];
            size_t nelem = 0;
            const char* err = NULL;
            void* buf = mcx_jdata_decode(Shapes, dims, 4, &ndim, type, sizeof(type), &nelem, &err);

            if (!buf) {
                MCX_ERROR(-1, err);
            }

            int eb = mcx_jdata_typebytes(type);

            if (strstr(type, "double") || strstr(type, "int64") || !strcmp(type, "single") || !strcmp(type, "float32")) {
                free(buf);
                MCX_ERROR(-1, "8-byte and floating-point JData volumes are not supported");
            }

            cfg->dim.x = dims[0];
            cfg->dim.y = dims[1];
            cfg->dim.z = dims[2];
            cfg->mediabyte = (unsigned int)eb;
            free(cfg->vol);
            cfg->vol = (unsigned int*)malloc((nelem ? nelem : 1) * sizeof(unsigned int));

            for (size_t i = 0; i < nelem; i++) {
                cfg->vol[i] = (eb == 1) ? ((unsigned char*)buf)[i] : (eb == 2) ? ((unsigned short*)buf)[i] :
                              ((unsigned int*)buf)[i];

                if (cfg->vol[i] >= cfg->medianum) {
                    free(buf);
                    MCX_ERROR(-6, "medium index exceeds the specified medium types");
                }
            }

            free(buf);
            cfg->isrowmajor = 1;   /* JData arrays are row-major */
        } else if (Shapes) {
            char err[256] = {0};
            int status = mcx_rasterize_shapes(root, &cfg->vol, &cfg->dim, cfg->issrcfrom0, cfg->isrowmajor, err);

            if (status) {
                MCX_ERROR(status, err);
            }
        } else {
            MCX_ERROR(-1, "You must either define Domain.VolumeFile, or define a Shapes section");
        }
    } else if (Shapes) {
        MCX_ERROR(-1, "You can not specify both Domain.VolumeFile and Shapes sections");
    }

    mcx_prepdomain(filename, cfg);
    cfg->his.maxmedia = cfg->medianum - 1;
    cfg->his.detnum = cfg->detnum;
    cfg->his.srcnum = cfg->srcnum;
    cfg->his.savedetflag = cfg->savedetflag;
    cfg->his.totalsource = ((cfg->srcid <= 0) ? cfg->extrasrclen + 1 : 1);
    return 0;
}

int mcx_loadjson_string(const char* json, Config* cfg) {
    char err[512] = {0};
    MJNode* root = mj_parse(json, err, sizeof(err));

    if (!root) {
        MCX_ERROR(-1, err[0] ? err : "invalid JSON input");
    }

    /* free the tree even when the loader errors out */
    jmp_buf jb;
    void* old = mcx_push_errjmp(&jb);
    int rc = setjmp(jb);

    if (rc == 0) {
        mcx_loadjson(root, cfg);
    }

    mcx_pop_errjmp(old);
    mj_free(root);

    if (rc) {
        char msg[2048];
        snprintf(msg, sizeof(msg), "%s", mcx_errmsg);

        if (mcx_errjmp) {
            snprintf(mcx_errmsg, sizeof(mcx_errmsg), "%s", msg);
            longjmp(*mcx_errjmp, rc);
        }

        fprintf(stdout, "\n%s\n", msg);
        exit(rc);
    }

    return 0;
}

/* ---------------- legacy .inp reader (reference src/mcx_utils.c:2087-2398) ---------------- */
static void skip_line(FILE* in, char* comment, size_t len, char** comm) {
    *comm = fgets(comment, (int)len, in);
}

void mcx_loadconfig(FILE* in, Config* cfg) {
    unsigned int gates, itmp;
    size_t count;
    float dtmp;
    char filename[MAX_PATH_LENGTH] = {0}, comment[MAX_FULL_PATH], *comm;

    MCX_ASSERT(fscanf(in, "%zu", &count) == 1);

    if (cfg->nphoton == 0) {
        cfg->nphoton = count;
    }

    skip_line(in, comment, sizeof(comment), &comm);

    if (cfg->seed == 0) {
        MCX_ASSERT(fscanf(in, "%d", &cfg->seed) == 1);
    } else {
        MCX_ASSERT(fscanf(in, "%u", &itmp) == 1);
    }

    skip_line(in, comment, sizeof(comment), &comm);
    MCX_ASSERT(fscanf(in, "%f %f %f", &cfg->srcpos.x, &cfg->srcpos.y, &cfg->srcpos.z) == 3);
    skip_line(in, comment, sizeof(comment), &comm);

    if (cfg->issrcfrom0 == 0 && comm && sscanf(comm, "%u", &itmp) == 1) {
        cfg->issrcfrom0 = (char)itmp;
    }

    if (!cfg->issrcfrom0) {
        cfg->srcpos.x--;
        cfg->srcpos.y--;
        cfg->srcpos.z--;
    }

    MCX_ASSERT(fscanf(in, "%f %f %f", &cfg->srcdir.x, &cfg->srcdir.y, &cfg->srcdir.z) == 3);
    skip_line(in, comment, sizeof(comment), &comm);

    if (comm && sscanf(comm, "%f", &dtmp) == 1) {
        cfg->srcdir.w = dtmp;
    }

    MCX_ASSERT(fscanf(in, "%f %f %f", &cfg->tstart, &cfg->tend, &cfg->tstep) == 3);
    skip_line(in, comment, sizeof(comment), &comm);

    if (cfg->tstart > cfg->tend || cfg->tstep == 0.f) {
        MCX_ERROR(-9, "incorrect time gate settings");
    }

    gates = (unsigned int)((cfg->tend - cfg->tstart) / cfg->tstep + 0.5);

    if (cfg->maxgate == 0 || cfg->maxgate > gates) {
        cfg->maxgate = gates;
    }

    MCX_ASSERT(fscanf(in, "%1023s", filename) == 1);

    if (cfg->rootpath[0]) {
        char full[MAX_FULL_PATH];
        snprintf(full, sizeof(full), "%s/%s", cfg->rootpath, filename);
        snprintf(filename, sizeof(filename), "%s", full);
    }

    skip_line(in, comment, sizeof(comment), &comm);
    MCX_ASSERT(fscanf(in, "%f %u %u %u", &cfg->steps.x, &cfg->dim.x, &cfg->crop0.x, &cfg->crop1.x) == 4);
    skip_line(in, comment, sizeof(comment), &comm);
    MCX_ASSERT(fscanf(in, "%f %u %u %u", &cfg->steps.y, &cfg->dim.y, &cfg->crop0.y, &cfg->crop1.y) == 4);
    skip_line(in, comment, sizeof(comment), &comm);
    MCX_ASSERT(fscanf(in, "%f %u %u %u", &cfg->steps.z, &cfg->dim.z, &cfg->crop0.z, &cfg->crop1.z) == 4);
    skip_line(in, comment, sizeof(comment), &comm);

    if (cfg->steps.x != cfg->steps.y || cfg->steps.y != cfg->steps.z) {
        MCX_ERROR(-9, "MCX currently does not support anisotropic voxels");
    }

    if (cfg->steps.x != 1.f && cfg->unitinmm == 1.f) {
        cfg->unitinmm = cfg->steps.x;
    }

    if (cfg->unitinmm != 1.f) {
        cfg->steps.x = cfg->steps.y = cfg->steps.z = cfg->unitinmm;
    }

    MCX_ASSERT(fscanf(in, "%u", &cfg->medianum) == 1);
    cfg->medianum++;
    skip_line(in, comment, sizeof(comment), &comm);
    free(cfg->prop);
    cfg->prop = (Medium*)calloc(cfg->medianum, sizeof(Medium));
    cfg->prop[0].g = 1.f;
    cfg->prop[0].n = 1.f;

    for (unsigned int i = 1; i < cfg->medianum; i++) {
        /* .inp media rows are mus, g, mua, n */
        MCX_ASSERT(fscanf(in, "%f %f %f %f", &cfg->prop[i].mus, &cfg->prop[i].g, &cfg->prop[i].mua, &cfg->prop[i].n) == 4);
        skip_line(in, comment, sizeof(comment), &comm);
    }

    MCX_ASSERT(fscanf(in, "%u %f", &cfg->detnum, &cfg->detradius) == 2);
    skip_line(in, comment, sizeof(comment), &comm);
    free(cfg->detpos);
    cfg->detpos = (MCXFloat4*)calloc(cfg->detnum ? cfg->detnum : 1, sizeof(MCXFloat4));

    if (cfg->issavedet && cfg->detnum == 0) {
        cfg->issavedet = 0;
    }

    for (unsigned int i = 0; i < cfg->detnum; i++) {
        MCX_ASSERT(fscanf(in, "%f %f %f", &cfg->detpos[i].x, &cfg->detpos[i].y, &cfg->detpos[i].z) == 3);
        cfg->detpos[i].w = cfg->detradius;

        if (!cfg->issrcfrom0) {
            cfg->detpos[i].x--;
            cfg->detpos[i].y--;
            cfg->detpos[i].z--;
        }

        skip_line(in, comment, sizeof(comment), &comm);

        if (comm && sscanf(comm, "%f", &dtmp) == 1) {
            cfg->detpos[i].w = dtmp;
        }
    }

    mcx_prepdomain(filename, cfg);
    cfg->his.maxmedia = cfg->medianum - 1;
    cfg->his.detnum = cfg->detnum;
    cfg->his.srcnum = cfg->srcnum;
    cfg->his.savedetflag = cfg->savedetflag;
    cfg->his.totalsource = ((cfg->srcid <= 0) ? cfg->extrasrclen + 1 : 1);

    char strtypestr[MAX_SESSION_LENGTH] = {0};

    if (fscanf(in, "%32s", strtypestr) == 1 && strtypestr[0]) {
        int st = mcx_keylookup(strtypestr, srctypeid);

        if (st == -1) {
            MCX_ERROR(-6, "the specified source type is not supported");
        }

        skip_line(in, comment, sizeof(comment), &comm);
        cfg->srctype = (char)st;
        MCX_ASSERT(fscanf(in, "%f %f %f %f", &cfg->srcparam1.x, &cfg->srcparam1.y, &cfg->srcparam1.z,
                          &cfg->srcparam1.w) == 4);
        MCX_ASSERT(fscanf(in, "%f %f %f %f", &cfg->srcparam2.x, &cfg->srcparam2.y, &cfg->srcparam2.z,
                          &cfg->srcparam2.w) == 4);
    }
}

/* .json or .inp input file (reference src/mcx_utils.c:2011-2070) */
void mcx_readconfig(char* fname, Config* cfg) {
    if (fname[0] == 0) {
        mcx_loadconfig(stdin, cfg);
        return;
    }

    FILE* fp = fopen(fname, "rt");

    if (!fp) {
        MCX_ERROR(-2, "can not load the specified config file");
    }

    if (strstr(fname, ".json") || strstr(fname, ".JSON")) {
        fseek(fp, 0, SEEK_END);
        long len = ftell(fp);
        rewind(fp);
        char* buf = (char*)malloc(len + 1);

        if (fread(buf, 1, len, fp) != (size_t)len) {
            fclose(fp);
            free(buf);
            MCX_ERROR(-2, "reading input file is terminated");
        }

        buf[len] = '\0';
        fclose(fp);

        if (cfg->rootpath[0] == '\0') {
            const char* slash = strrchr(fname, '/');

            if (slash) {
                size_t n = (size_t)(slash - fname);
                n = MIN(n, (size_t)MAX_PATH_LENGTH - 1);
                memcpy(cfg->rootpath, fname, n);
                cfg->rootpath[n] = '\0';
            }
        }

        jmp_buf jb;
        void* old = mcx_push_errjmp(&jb);
        int rc = setjmp(jb);

        if (rc == 0) {
            mcx_loadjson_string(buf, cfg);
        }

        mcx_pop_errjmp(old);
        free(buf);

        if (rc) {
            char msg[2048];
            snprintf(msg, sizeof(msg), "%s", mcx_errmsg);

            if (mcx_errjmp) {
                snprintf(mcx_errmsg, sizeof(mcx_errmsg), "%s", msg);
                longjmp(*mcx_errjmp, rc);
            }

            fprintf(stdout, "\n%s\n", msg);
            exit(rc);
        }
    } else {
        mcx_loadconfig(fp, cfg);
        fclose(fp);
    }
}

/* reference src/mcx_utils.c:3910-3958 */
void mcx_prepdomain(char* filename, Config* cfg) {
    if (filename && filename[0]) {
        if (!cfg->vol) {
            if (strstr(filename, ".json")) {
                FILE* fp = fopen(filename, "rb");

                if (!fp) {
                    MCX_ERROR(-2, "Can not read the JSON shape file");
                }

                fseek(fp, 0, SEEK_END);
                long len = ftell(fp);
                rewind(fp);
                char* buf = (char*)malloc(len + 1);
                size_t got = fread(buf, 1, len, fp);
                buf[got] = '\0';
                fclose(fp);
                int rc = mcx_try_parse_shapestring(cfg, buf);
                free(buf);

                if (rc) {
                    MCX_ERROR(rc, mcx_errmsg);
                }
            } else {
                mcx_loadvolume(filename, cfg);
            }
        }
    } else if (!cfg->vol) {
        MCX_ERROR(-4, "one must specify a binary volume file in order to run the simulation");
    }

    if (cfg->seed == SEED_FROM_FILE && cfg->seedfile[0]) {
        if (strstr(cfg->seedfile, ".jdat")) {
            mcx_loadseedjdat(cfg->seedfile, cfg);
        } else {
            mcx_loadseedfile(cfg);
        }
    }

    mcx_preprocess(cfg);
}

int mcx_parse_shapestring(Config* cfg, const char* shapedata) {
    char err[512] = {0};
    MJNode* root = mj_parse(shapedata, err, sizeof(err));

    if (!root) {
        MCX_ERROR(-2, err[0] ? err : "Invalid JSON shape data");
    }

    char rerr[256] = {0};
    int status = mcx_rasterize_shapes(root, &cfg->vol, &cfg->dim, cfg->issrcfrom0, cfg->isrowmajor, rerr);
    mj_free(root);

    if (status) {
        MCX_ERROR(status, rerr);
    }

    return 0;
}

/* ---------------- detector mask (reference src/mcx_utils.c:4085-4198) ---------------- */
/* Marks (bit 31) every non-zero voxel that touches a zero voxel or the domain
   boundary (26-neighbourhood) and lies within R+1/2 of a detector centre, so
   the kernel only searches detectors for photons escaping through such voxels */
void mcx_maskdet(Config* cfg) {
    const int nx = (int)cfg->dim.x, ny = (int)cfg->dim.y, nz = (int)cfg->dim.z;
    const int px = nx + 2, py = ny + 2, pz = nz + 2;
    int isonecube = (nx == 1 && ny == 1 && nz == 1);
    unsigned int* pad = (unsigned int*)calloc((size_t)px * py * pz, sizeof(unsigned int));
    static const float corner[8][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}, {1, 1, 0}, {1, 0, 1}, {0, 1, 1},
        {1, 1, 1}
    };

    for (int z = 0; z < nz; z++)
        for (int y = 0; y < ny; y++)
            for (int x = 0; x < nx; x++) {
                pad[((size_t)(z + 1) * py + (y + 1)) * px + (x + 1)] = cfg->vol[((size_t)z * ny + y) * nx + x];
            }

    for (unsigned int d = 0; d < cfg->detnum; d++) {
        const MCXFloat4 det = cfg->detpos[d];
        float reach = (det.w + 1.7321f) * (det.w + 1.7321f);
        float search = (det.w + 1.f) * (det.w + 1.f);
        float keep = (det.w + 0.5f * (1.f + isonecube)) * (det.w + 0.5f * (1.f + isonecube));
        unsigned int count = 0;

        for (float z = -det.w - 1.f; z <= det.w + 1.f; z += 0.5f) {
            float iz = z + det.z;

            for (float y = -det.w - 1.f; y <= det.w + 1.f; y += 0.5f) {
                float iy = y + det.y;

                for (float x = -det.w - 1.f; x <= det.w + 1.f; x += 0.5f) {
                    float ix = x + det.x;

                    if (iz < 0 || ix < 0 || iy < 0 || ix >= nx || iy >= ny || iz >= nz ||
                            x * x + y * y + z * z > search) {
                        continue;
                    }

                    float mind2 = VERY_BIG;

                    for (int c = 0; c < 8; c++) {
                        float rx = (int)ix - det.x + corner[c][0];
                        float ry = (int)iy - det.y + corner[c][1];
                        float rz = (int)iz - det.z + corner[c][2];
                        float d2 = rx * rx + ry * ry + rz * rz;

                        if (d2 > reach) {
                            mind2 = VERY_BIG;
                            break;
                        }

                        if (d2 < mind2) {
                            mind2 = d2;
                        }
                    }

                    if (mind2 == VERY_BIG || mind2 >= keep) {
                        continue;
                    }

                    size_t p = ((size_t)((int)(iz + 1.f)) * py + (int)(iy + 1.f)) * px + (int)(ix + 1.f);

                    if (cfg->mediabyte == MEDIA_2LABEL_SPLIT) {
                        /* src/mcx_utils.c:4157-4166: a split voxel whose lower side is background */
                        if (!(pad[p] >> 24) && ((pad[p] >> 16) & 0xFFu)) {
                            cfg->vol[((size_t)(int)iz * ny + (int)iy) * nx + (int)ix] |= DET_MASK;
                            count++;
                        }

                        continue;
                    }

                    if (!pad[p]) {
                        continue;
                    }

                    int interior = 1;

                    for (int dz = -1; dz <= 1 && interior; dz++)
                        for (int dy = -1; dy <= 1 && interior; dy++)
                            for (int dx = -1; dx <= 1; dx++) {
                                if (!pad[p + ((long)dz * py + dy) * px + dx]) {
                                    interior = 0;
                                    break;
                                }
                            }

                    if (!interior) {
                        cfg->vol[((size_t)(int)iz * ny + (int)iy) * nx + (int)ix] |= DET_MASK;
                        count++;
                    }
                }
            }
        }

        if (cfg->issavedet && count == 0) {
            fprintf(stderr, "MCX WARNING: detector %u is not located on an interface, please check coordinates.\n", d + 1);
        }
    }

    free(pad);
}

/* ---------------- preprocessing (reference src/mcx_utils.c:1521-1809) ---------------- */
/* scattering of the polarised media (src/mcx_utils.c:1483-1519): medium i+1 is
   a suspension of spheres (radius r in um, n_sph) in a background (n_med, mua)
   at wavelength lambda (nm).  Each gets its Mie table over NANGLES angles
   (theta_k = k degrees) in cfg->smatrix, g from the series, and mus from the
   particle density rho (1/um^3): mus = Qsca pi r^2 rho (1/mm).  A positive mus
   given for the medium instead sets rho so that the medium keeps that mus, or
   its reduced scattering mus (1 - g) when its g is below 1. */
void mcx_prep_polarized(Config* cfg) {
    double* mu = (double*)malloc(MCX_NANGLES * sizeof(double));

    for (int i = 0; i < MCX_NANGLES; i++) {
        mu[i] = cos(M_PI * i / (MCX_NANGLES - 1));
    }

    free(cfg->smatrix);
    cfg->smatrix = (MCXFloat4*)malloc((size_t)cfg->polmedianum * MCX_NANGLES * sizeof(MCXFloat4));

    for (unsigned int i = 0; i < cfg->polmedianum; i++) {
        POLMedium* pm = cfg->polprop + i;
        Medium* md = cfg->prop + i + 1;
        double qsca = 0.0, g = 0.0;
        md->mua = pm->mua;
        md->n = pm->nmed;
        const double x = 2.0 * M_PI * pm->r * pm->nmed / (cfg->lambda * 1e-3);
        const double area = M_PI * pm->r * pm->r;

        if (mcx_mie(x, pm->nsph / pm->nmed, MCX_NANGLES, mu, (float*)(cfg->smatrix + (size_t)i * MCX_NANGLES), &qsca,
                    &g)) {
            free(mu);
            MCX_ERROR(-6, "sphere size parameter must lie in (0, 20000]");
        }

        if (md->mus > EPS) {
            float target = md->mus;

            if (md->g < 1.f - EPS) {
                target = (float)(md->mus * (1.0f - md->g) / (1.0 - g));
            }

            pm->rho = (float)(target / qsca / area * 1e-3);
        }

        md->mus = (float)(qsca * area * pm->rho * 1e3);
        md->g = (float)g;
    }

    free(mu);
}

void mcx_preprocess(Config* cfg) {
    int isbcdet = 0;
    double len = sqrt((double)cfg->srcdir.x * cfg->srcdir.x + (double)cfg->srcdir.y * cfg->srcdir.y +
                      (double)cfg->srcdir.z * cfg->srcdir.z);

    if (len < EPS) {
        MCX_ERROR(-4, "source initial direction vector can not have a length of 0");
    }

    float inv = (float)(1.0 / len);
    cfg->srcdir.x *= inv;
    cfg->srcdir.y *= inv;
    cfg->srcdir.z *= inv;

    if (cfg->vol && cfg->polprop) {
        if (!(cfg->mediabyte <= 4)) {
            MCX_ERROR(-1, "Unsupported media format for polarized simulation");
        }

        if (cfg->medianum != cfg->polmedianum + 1) {
            MCX_ERROR(-6, "number of particle types does not match number of media");
        }

        if (cfg->lambda == 0.f) {
            MCX_ERROR(-1, "you must specify light wavelength lambda to run polarized photon simulation");
        }

        mcx_prep_polarized(cfg);
    }

    if (cfg->debuglevel & MCX_DEBUG_MOVE_ONLY) {
        cfg->issave2pt = 0;
        cfg->issavedet = 0;
    }

    if (cfg->debuglevel & MCX_DEBUG_RNG) {
        cfg->isnormalized = 0;
        cfg->issavedet = 0;
    }

    for (int i = 0; i < 6; i++)
        if (cfg->bc[i] && cfg->bc[i] > 4 && mcx_lookupindex(cfg->bc + i, boundarycond)) {
            MCX_ERROR(-4, "unknown boundary condition specifier");
        }

    for (int i = 6; i < 12; i++) {
        if (cfg->bc[i] && cfg->bc[i] > 1 && mcx_lookupindex(cfg->bc + i, boundarydetflag)) {
            MCX_ERROR(-4, "unknown boundary detection flags");
        }

        if (cfg->bc[i]) {
            isbcdet = 1;
        }
    }

    if (cfg->medianum == 0) {
        MCX_ERROR(-4, "you must define the 'prop' field in the input structure");
    }

    if (cfg->dim.x == 0 || cfg->dim.y == 0 || cfg->dim.z == 0) {
        MCX_ERROR(-4, "the 'vol' field in the input structure can not be empty");
    }

    if ((cfg->srctype == MCX_SRC_PATTERN || cfg->srctype == MCX_SRC_PATTERN3D) && cfg->srcpattern == NULL) {
        MCX_ERROR(-4, "the 'srcpattern' field can not be empty when your 'srctype' is 'pattern'");
    }

    /* the pixel index of a pattern launch is (int)(r * 0.9998 * n) per axis with
       the row length (int)n (src/mcx_core.cl:1678-1727): it stays inside the
       pattern only for whole, positive pixel counts */
    if (cfg->srctype == MCX_SRC_PATTERN || cfg->srctype == MCX_SRC_PATTERN3D) {
        const float dims[3] = {cfg->srctype == MCX_SRC_PATTERN ? cfg->srcparam1.w : cfg->srcparam1.x,
                               cfg->srctype == MCX_SRC_PATTERN ? cfg->srcparam2.w : cfg->srcparam1.y,
                               cfg->srctype == MCX_SRC_PATTERN ? 1.f : cfg->srcparam1.z
                              };

        for (int i = 0; i < 3; i++) {
            if (!(dims[i] >= 1.f) || dims[i] != floorf(dims[i]) || dims[i] > 65536.f) {
                MCX_ERROR(-4, "pattern sources need whole, positive pixel counts (srcparam1.w/srcparam2.w for "
                          "'pattern', srcparam1.x/y/z for 'pattern3d')");
            }
        }
    }

    if (cfg->steps.x != 1.f && cfg->unitinmm == 1.f) {
        cfg->unitinmm = cfg->steps.x;
    }

    if (cfg->unitinmm != 1.f) {
        cfg->steps.x = cfg->steps.y = cfg->steps.z = cfg->unitinmm;

        for (unsigned int i = 1; i < cfg->medianum; i++) {
            cfg->prop[i].mus *= cfg->unitinmm;
            cfg->prop[i].mua *= cfg->unitinmm;
        }
    }

    if (cfg->isrowmajor) {
        if (cfg->mediabyte == MEDIA_2LABEL_SPLIT || cfg->mediabyte == MEDIA_MUA_MUS_FLOAT) {
            mcx_convertrow2col64(&cfg->vol, &cfg->dim);
        } else {
            mcx_convertrow2col(&cfg->vol, &cfg->dim);
        }

        cfg->isrowmajor = 0;
    }

    if (cfg->issavedet && cfg->detnum == 0 && isbcdet == 0) {
        cfg->issavedet = 0;
    }

    if (cfg->issavedet == 0) {
        cfg->issaveexit = 0;
        cfg->ismomentum = 0;

        if (cfg->seed != SEED_FROM_FILE) {
            cfg->savedetflag = 0;
        }
    }

    if (cfg->respin == 0) {
        MCX_ERROR(-1, "respin number can not be 0, check your -r/--repeat input or cfg.respin value");
    }

    if (cfg->seed == SEED_FROM_FILE && cfg->seedfile[0]) {
        if (cfg->respin > 1 || cfg->respin < 0) {
            cfg->respin = 1;
            fprintf(stderr, "WARNING: respin is disabled in the replay mode\n");
        }
    }

    if (cfg->replaydet > (int)cfg->detnum) {
        MCX_ERROR(-4, "replay detector ID exceeds the maximum detector number");
    }

    if (cfg->replaydet == -1 && cfg->detnum == 1) {
        cfg->replaydet = 1;
    }

    for (unsigned int i = 0; i < cfg->medianum; i++) {
        if (cfg->prop[i].mus == 0.f && !cfg->polmedianum) {
            cfg->prop[i].mus = EPS;
        }
    }

    if (cfg->srcpos.w == 0.f) {
        cfg->srcpos.w = 1.f;
    }

    if (cfg->extrasrclen) {
        if (cfg->srcnum > 1) {
            MCX_ERROR(-4, "simulating multiple sources currently can not be used with photon-sharing");
        }

        if (cfg->srcid > (int)cfg->extrasrclen + 1) {
            MCX_ERROR(-4, "srcid exceeds total defined source count");
        }

        for (unsigned int i = 0; i < cfg->extrasrclen; i++) {
            if (cfg->srcdata[i].srcpos.w == 0.f) {
                cfg->srcdata[i].srcpos.w = 1.f;
            }
        }
    }

    if (cfg->vol) {
        size_t n = (size_t)cfg->dim.x * cfg->dim.y * cfg->dim.z;

        if (cfg->mediabyte <= 4) {
            unsigned int maxlabel = 0;

            for (size_t i = 0; i < n; i++) {
                maxlabel = MAX(maxlabel, (cfg->vol[i] & MED_MASK));
            }

            if (cfg->medianum <= maxlabel) {
                MCX_ERROR(-4, "input media optical properties are less than the labels in the volume");
            }
        } else if (cfg->mediabyte == MEDIA_LABEL_HALF) {
            /* the label in the low 14 bits indexes the media table (the reference
               reads past the table for a larger one, src/mcx_core.cl:1145) */
            for (size_t i = 0; i < n; i++) {
                if ((cfg->vol[i] & MED_MASK) && (cfg->vol[i] & 0x3FFFu) >= cfg->medianum) {
                    MCX_ERROR(-4, "input media optical properties are less than the labels in the volume");
                }
            }
        } else if (cfg->mediabyte == MEDIA_2LABEL_SPLIT) {
            /* both labels index the media table, and the lower one must leave bit 31
               (the detector flag) clear: the kernel masks it off (the reference reads
               past the table for larger labels) */
            const unsigned char* b = (const unsigned char*)cfg->vol;

            for (size_t i = 0; i < n; i++) {
                if (b[i << 3] >= cfg->medianum || b[(i << 3) + 1] >= cfg->medianum || b[i << 3] >= 128) {
                    MCX_ERROR(-4, "svmc labels must be below the number of media (and the lower label below 128)");
                }
            }

            mcx_split_svmc(cfg);
        } else if (cfg->mediabyte != MEDIA_MUA_MUS_FLOAT && cfg->mediabyte != MEDIA_MUA_FLOAT &&
                   cfg->mediabyte != MEDIA_AS_F2H && cfg->mediabyte != MEDIA_AS_HALF &&
                   cfg->mediabyte != MEDIA_ASGN_BYTE && cfg->mediabyte != MEDIA_AS_SHORT) {
            MCX_ERROR(-4, "this build supports the label (byte/short/integer), svmc, labelplus, muamus_float, "
                      "mua_float, muamus_half, asgn_byte, muamus_short and muamus_f32 media formats");
        }
    }

    /* src/mcx_utils.c:1760-1766 */
    if ((cfg->mediabyte == MEDIA_AS_F2H || cfg->mediabyte == MEDIA_MUA_FLOAT || cfg->mediabyte == MEDIA_AS_HALF) &&
            cfg->medianum < 2) {
        MCX_ERROR(-4, "the 'prop' field must contain at least 2 rows for the requested media format");
    }

    if ((cfg->mediabyte == MEDIA_ASGN_BYTE || cfg->mediabyte == MEDIA_AS_SHORT) && cfg->medianum < 3) {
        MCX_ERROR(-4, "the 'prop' field must contain at least 3 rows for the requested media format");
    }

    /* point sources: pre-compute the launch voxel and its label, stored as raw
       bits in srcparam2.z/.w (reference src/mcx_utils.c:1718-1749) */
    if (cfg->srctype <= MCX_SRC_CONE || cfg->srctype == MCX_SRC_ARCSINE || cfg->srctype == MCX_SRC_ZGAUSSIAN) {
        union {
            float f;
            unsigned int u;
        } f2i;

        if (cfg->srcpos.x < 0.f || cfg->srcpos.y < 0.f || cfg->srcpos.z < 0.f || cfg->srcpos.x >= cfg->dim.x ||
                cfg->srcpos.y >= cfg->dim.y || cfg->srcpos.z >= cfg->dim.z) {
            f2i.u = 0;
            cfg->srcparam2.z = f2i.f;
            cfg->srcparam2.w = f2i.f;
        } else {
            f2i.u = (unsigned int)((int)floorf(cfg->srcpos.z) * (cfg->dim.y * cfg->dim.x) +
                                   (int)floorf(cfg->srcpos.y) * cfg->dim.x + (int)floorf(cfg->srcpos.x));
            cfg->srcparam2.z = f2i.f;
            unsigned int lab = (cfg->mediabyte == MEDIA_MUA_MUS_FLOAT) ? 1u : (cfg->vol[f2i.u] & MED_MASK);
            f2i.u = lab;
            cfg->srcparam2.w = f2i.f;
        }

        for (unsigned int i = 0; i < cfg->extrasrclen; i++) {
            ExtraSrc* s = cfg->srcdata + i;

            if (s->srcpos.x < 0.f || s->srcpos.y < 0.f || s->srcpos.z < 0.f || s->srcpos.x >= cfg->dim.x ||
                    s->srcpos.y >= cfg->dim.y || s->srcpos.z >= cfg->dim.z) {
                f2i.u = 0;
                s->srcparam2.z = f2i.f;
                s->srcparam2.w = f2i.f;
            } else {
                f2i.u = (unsigned int)((int)floorf(s->srcpos.z) * (cfg->dim.y * cfg->dim.x) +
                                       (int)floorf(s->srcpos.y) * cfg->dim.x + (int)floorf(s->srcpos.x));
                s->srcparam2.z = f2i.f;
                unsigned int lab = (cfg->mediabyte == MEDIA_MUA_MUS_FLOAT) ? 1u : (cfg->vol[f2i.u] & MED_MASK);
                f2i.u = lab;
                s->srcparam2.w = f2i.f;
            }
        }
    }

    /* the hyperboloid launch keeps srcparam2.z/.w as its launch voxel and label
       (src/mcx_core.cl:1820-1852); they are the zero defaults, so the launch
       finds its voxel by skipvoid.  Anything else there would index outside the
       volume or the medium table, so it is cleared. */
    if (cfg->srctype == MCX_SRC_HYPERBOLOID_GAUSSIAN) {
        cfg->srcparam2.z = 0.f;
        cfg->srcparam2.w = 0.f;
    }

    /* one word per voxel: the detector mask marks non-zero words; split voxels:
       those holding background below a labelled upper side (src/mcx_utils.c:1751-1753) */
    if (cfg->issavedet && (cfg->mediabyte <= 4 || cfg->mediabyte == MEDIA_2LABEL_SPLIT ||
                           (cfg->mediabyte >= MEDIA_LABEL_HALF && cfg->mediabyte <= MEDIA_AS_SHORT))) {
        mcx_maskdet(cfg);
    }

    for (int i = 0; i < MAX_DEVICE; i++)
        if (cfg->deviceid[i] == '0') {
            cfg->deviceid[i] = '\0';
        }

    if (cfg->ismomentum) {
        cfg->savedetflag |= 0x8;
    }

    if (cfg->issaveexit) {
        cfg->savedetflag |= 0x10 | 0x20;
    }

    /* polarised runs record partial paths, exit directions, initial weights and
       the Stokes vector of each detected photon (src/mcx_utils.c:1777-1784) */
    if (cfg->polmedianum) {
        cfg->savedetflag |= 0x4u | 0x20u | 0x40u | 0x80u;
    } else {
        cfg->savedetflag &= ~0x80u;
    }

    if (cfg->issavedet && cfg->savedetflag == 0) {
        cfg->savedetflag = 0x5;
    }

    if (cfg->mediabyte >= 100 && cfg->savedetflag) {
        cfg->savedetflag &= ~(0x2u | 0x4u | 0x8u);
    }

    /* exit tally per voxel (src/mcx_utils.c:1797-1806): the detected-photon
       buffer holds one {weight, partial paths} record per voxel */
    if (cfg->issaveref > 1) {
        if (cfg->issavedet == 0) {
            MCX_ERROR(-4, "you must have at least two outputs if issaveref is greater than 1");
        }

        if ((size_t)cfg->dim.x * cfg->dim.y * cfg->dim.z > cfg->maxdetphoton) {
            cfg->maxdetphoton = cfg->dim.x * cfg->dim.y * cfg->dim.z;
        }

        cfg->savedetflag = 0x5;
    }
}

/* ---------------- replay preparation (reference src/mcx_utils.c:1355-1421) ---------------- */
void mcx_replayinit(Config* cfg, float* detps, int dimdetps[2], int seedbyte) {
    if (cfg->seed == SEED_FROM_FILE && detps == NULL) {
        MCX_ERROR(-6, "you give cfg.seed for replay, but did not specify cfg.detphotons.\nPlease define it as the detphoton output from the baseline simulation\n");
    }

    if (detps == NULL || cfg->seed != SEED_FROM_FILE) {
        return;
    }

    if (cfg->nphoton != (size_t)dimdetps[1]) {
        MCX_ERROR(-6, "the column numbers of detphotons and seed do not match\n");
    }

    if (seedbyte == 0) {
        MCX_ERROR(-6, "the seed input is empty");
    }

    int hasdetid = SAVE_DETID(cfg->savedetflag);
    int offset = SAVE_NSCAT(cfg->savedetflag) * (cfg->medianum - 1);

    if (((!hasdetid) && cfg->detnum > 1) || !SAVE_PPATH(cfg->savedetflag)) {
        MCX_ERROR(-6, "please rerun the baseline simulation and save detector ID (D) and partial-path (P) using cfg.savedetflag='dp' ");
    }

    free(cfg->replay.weight);
    free(cfg->replay.tof);
    free(cfg->replay.detid);
    cfg->replay.weight = (float*)malloc(cfg->nphoton * sizeof(float));
    cfg->replay.tof = (float*)calloc(cfg->nphoton, sizeof(float));
    cfg->replay.detid = (int*)calloc(cfg->nphoton, sizeof(int));
    size_t kept = 0;

    for (int i = 0; i < dimdetps[1]; i++) {
        const float* rec = detps + (size_t)i * dimdetps[0];
        int detword = (int)rec[0];

        if (cfg->replaydet > 0 && cfg->replaydet != (detword & 0xFFFF)) {
            continue;
        }

        if (cfg->srcid > 0 && !((detword & 0xFFFF0000) && cfg->srcid == ((detword & 0xFFFF0000) >> 16))) {
            continue;
        }

        if ((size_t)i != kept) {
            memcpy((char*)cfg->replay.seed + kept * seedbyte, (char*)cfg->replay.seed + (size_t)i * seedbyte, seedbyte);
        }

        float w = 1.f, tof = 0.f;

        for (int j = hasdetid; j < (int)cfg->medianum - 1 + hasdetid; j++) {
            float plen = rec[offset + j];
            w *= expf(-cfg->prop[j - hasdetid + 1].mua * plen);
            plen *= cfg->unitinmm;
            tof += plen * R_C0 * cfg->prop[j - hasdetid + 1].n;
        }

        if (cfg->replayinweight) {
            w = cfg->replayinweight[i];   /* per-voxel media: the recorded exit weight */
        }

        cfg->replay.weight[kept] = w;
        cfg->replay.tof[kept] = tof;
        cfg->replay.detid[kept] = hasdetid ? detword : 1;

        if (tof < cfg->tstart || tof > cfg->tend) {
            continue;
        }

        kept++;
    }

    cfg->nphoton = kept;
}

/* ---------------- replay from files (reference src/mcx_utils.c:1432-1466, :3707-3897) ---------------- */
/* keep the photons of cfg->replaydet (all when <= 0) whose time of flight lies in
   [tstart, tend], compacting the seeds; ppath holds his->maxmedia path lengths
   per photon in grid units and the media are still in 1/mm here */
void mcx_replayprep(int* detid, float* ppath, History* his, Config* cfg) {
    size_t kept = 0;

    for (size_t i = 0; i < his->savedphoton; i++) {
        if (cfg->replaydet > 0 && !(detid && cfg->replaydet == (detid[i] & 0xFFFF))) {
            continue;
        }

        if (i != kept) {
            memmove((char*)cfg->replay.seed + kept * his->seedbyte, (char*)cfg->replay.seed + i * his->seedbyte,
                    his->seedbyte);
        }

        float w = 1.f, tof = 0.f;

        for (unsigned int j = 0; j < his->maxmedia; j++) {
            float plen = ppath[i * his->maxmedia + j] * his->unitinmm;
            w *= expf(-cfg->prop[j + 1].mua * plen);
            tof += plen * R_C0 * cfg->prop[j + 1].n;
        }

        cfg->replay.weight[kept] = w;
        cfg->replay.tof[kept] = tof;
        cfg->replay.detid[kept] = detid ? detid[i] : 1;

        if (tof < cfg->tstart || tof > cfg->tend) {
            continue;
        }

        kept++;
    }

    cfg->nphoton = kept;
    cfg->minenergy = 0.f;
}

static void replay_alloc(Config* cfg, size_t n) {
    free(cfg->replay.weight);
    free(cfg->replay.tof);
    free(cfg->replay.detid);
    cfg->replay.weight = (float*)malloc((n ? n : 1) * sizeof(float));
    cfg->replay.tof = (float*)calloc(n ? n : 1, sizeof(float));
    cfg->replay.detid = (int*)calloc(n ? n : 1, sizeof(int));

    if (!cfg->replay.weight || !cfg->replay.tof || !cfg->replay.detid) {
        MCX_ERROR(-7, "can not allocate memory");
    }
}

/* reference mcx_loadseedfile (src/mcx_utils.c:3806-3897): seeds of an .mch file
   from a run with -q 1; replay outputs (j, p, r) also take the weights, times of
   flight and detector ids from its D and P columns */
void mcx_loadseedfile(Config* cfg) {
    History his;
    FILE* fp = fopen(cfg->seedfile, "rb");

    if (!fp) {
        MCX_ERROR(-7, "can not open the specified history file");
    }

    if (fread(&his, sizeof(History), 1, fp) != 1) {
        fclose(fp);
        MCX_ERROR(-7, "error when reading the history file");
    }

    if (his.savedphoton == 0 || his.seedbyte == 0) {
        fclose(fp);
        MCX_ERROR(-7, "history file does not contain seed data, please re-run your simulation with '-q 1'");
    }

    if (his.maxmedia != cfg->medianum - 1) {
        fclose(fp);
        MCX_ERROR(-7, "the history file was generated with a different media setting");
    }

    if (his.seedbyte != 16) {
        fclose(fp);
        MCX_ERROR(-7, "the history file holds seeds of another random number generator");
    }

    size_t nrec = his.savedphoton;
    size_t nval = nrec * his.colcount;
    float* rec = (float*)malloc((nval > 0 ? nval : 1) * sizeof(float));
    free(cfg->replay.seed);
    cfg->replay.seed = malloc(nrec * his.seedbyte);

    if (!rec || !cfg->replay.seed) {
        free(rec);
        fclose(fp);
        MCX_ERROR(-7, "can not allocate memory");
    }

    if (fread(rec, sizeof(float) * his.colcount, nrec, fp) != nrec ||
            fread(cfg->replay.seed, his.seedbyte, nrec, fp) != nrec) {
        free(rec);
        fclose(fp);
        MCX_ERROR(-7, "error when reading the seed data");
    }

    fclose(fp);
    cfg->seed = SEED_FROM_FILE;
    cfg->nphoton = nrec;
    replay_alloc(cfg, nrec);

    int hasdetid = SAVE_DETID(his.savedetflag);
    int isreplayout = (cfg->outputtype == otJacobian || cfg->outputtype == otWP || cfg->outputtype == otDCS ||
                       cfg->outputtype == otRF || cfg->outputtype == otRFmus || cfg->outputtype == otWLTOF ||
                       cfg->outputtype == otWPTOF);

    if (!isreplayout) {
        /* a plain replay relaunches every saved photon with its own weight */
        for (size_t i = 0; i < nrec; i++) {
            cfg->replay.weight[i] = 1.f;
            cfg->replay.detid[i] = hasdetid ? (int)rec[i * his.colcount] : 1;
        }

        free(rec);
        return;
    }

    if ((!hasdetid && cfg->detnum > 1) || !SAVE_PPATH(his.savedetflag)) {
        free(rec);
        MCX_ERROR(-7, "please rerun the baseline simulation and save detector ID (D) and partial-path (P) using '-w DP'");
    }

    /* gather the detector ids and the partial paths into mcx_replayprep's layout */
    unsigned int off = hasdetid + SAVE_NSCAT(his.savedetflag) * his.maxmedia;
    int* detid = (int*)malloc(nrec * sizeof(int));
    size_t npath = nrec * his.maxmedia;
    float* pp = (float*)malloc((npath > 0 ? npath : 1) * sizeof(float));

    if (!detid || !pp) {
        free(detid);
        free(pp);
        free(rec);
        MCX_ERROR(-7, "can not allocate memory");
    }

    for (size_t i = 0; i < nrec; i++) {
        detid[i] = hasdetid ? (int)rec[i * his.colcount] : 1;
        memcpy(pp + i * his.maxmedia, rec + i * his.colcount + off, his.maxmedia * sizeof(float));
    }

    mcx_replayprep(detid, pp, &his, cfg);
    free(detid);
    free(pp);
    free(rec);
}

/* reference mcx_loadseedjdat (src/mcx_utils.c:3707-3800): MCXData.PhotonData
   .seed, .detid and .ppath of a _detp.jdat file */
void mcx_loadseedjdat(char* filename, Config* cfg) {
    FILE* fp = fopen(filename, "rb");

    if (!fp) {
        MCX_ERROR(-6, "fail to open the specified seed jdata file");
    }

    fseek(fp, 0, SEEK_END);
    long len = ftell(fp);
    rewind(fp);
    char* text = (char*)malloc((size_t)len + 1);

    if (!text || fread(text, 1, (size_t)len, fp) != (size_t)len) {
        free(text);
        fclose(fp);
        MCX_ERROR(-2, "reading input file is terminated");
    }

    fclose(fp);
    text[len] = '\0';
    char perr[256] = {0};
    MJNode* root = mj_parse(text, perr, sizeof(perr));
    free(text);

    if (!root) {
        MCX_ERROR(-1, "invalid jdat file");
    }

    MJNode* data = mj_get(root, "MCXData");
    MJNode* info = data ? mj_get(data, "Info") : NULL;
    MJNode* photon = data ? mj_get(data, "PhotonData") : NULL;
    MJNode* jseed = photon ? mj_get(photon, "seed") : NULL;
    MJNode* jdet = photon ? mj_get(photon, "detid") : NULL;
    MJNode* jppath = photon ? mj_get(photon, "ppath") : NULL;

    if (!jseed || !jppath || !jdet || !info) {
        mj_free(root);
        MCX_ERROR(-1, "invalid jdat file, expect MCXData.PhotonData.seed, .detid and .ppath");
    }

    unsigned int sdims[2] = {0, 0}, pdims[2] = {0, 0}, ddims[2] = {0, 0};
    int ndim = 0;
    char type[32];
    size_t ns = 0, np = 0, nd = 0;
    const char* err = NULL;
    History his;
    memset(&his, 0, sizeof(his));
    unsigned char* seeds = (unsigned char*)mcx_jdata_decode(jseed, sdims, 2, &ndim, type, sizeof(type), &ns, &err);
    float* pp = seeds ? (float*)mcx_jdata_decode(jppath, pdims, 2, &ndim, type, sizeof(type), &np, &err) : NULL;
    int* detid = pp ? (int*)mcx_jdata_decode(jdet, ddims, 2, &ndim, type, sizeof(type), &nd, &err) : NULL;
    his.unitinmm = (float)mj_num(mj_get(info, "LengthUnit"), 1.0);
    mj_free(root);

    if (!detid || sdims[1] != 16 || pdims[0] != sdims[0] || ddims[0] != sdims[0] ||
            pdims[1] != cfg->medianum - 1) {
        free(seeds);
        free(pp);
        free(detid);
        MCX_ERROR(-1, err ? err : "the jdat seed, detid and ppath arrays do not match this simulation");
    }

    his.savedphoton = sdims[0];
    his.seedbyte = sdims[1];
    his.maxmedia = pdims[1];
    free(cfg->replay.seed);
    cfg->replay.seed = seeds;
    cfg->seed = SEED_FROM_FILE;
    cfg->nphoton = his.savedphoton;
    replay_alloc(cfg, his.savedphoton);
    mcx_replayprep(detid, pp, &his, cfg);
    free(pp);
    free(detid);
}

/* ---------------- validation (reference src/mcx_utils.c:1822-1964) ---------------- */
/* append the detectors as extra sources of the adjoint outputs: position and
   direction of each detector, unit weight, disk radius detpos.w in srcparam1.x;
   all sources then launch in one run with one field slot each (srcid -1) */
void mcx_append_detector_sources(Config* cfg) {
    const unsigned int n0 = cfg->extrasrclen;
    ExtraSrc* sd = (ExtraSrc*)realloc(cfg->srcdata, (n0 + cfg->detnum) * sizeof(ExtraSrc));

    if (!sd) {
        MCX_ERROR(-1, "cannot allocate the detector sources");
    }

    cfg->srcdata = sd;
    memset(sd + n0, 0, cfg->detnum * sizeof(ExtraSrc));

    for (unsigned int i = 0; i < cfg->detnum; i++) {
        sd[n0 + i].srcpos.x = cfg->detpos[i].x;
        sd[n0 + i].srcpos.y = cfg->detpos[i].y;
        sd[n0 + i].srcpos.z = cfg->detpos[i].z;
        sd[n0 + i].srcpos.w = 1.f;
        memcpy(&sd[n0 + i].srcdir, cfg->detdir + i, sizeof(sd[n0 + i].srcdir));
        sd[n0 + i].srcparam1.x = cfg->detpos[i].w;
    }

    cfg->extrasrclen += cfg->detnum;

    if (MCX_IS_ADJOINT_TYPE(cfg->outputtype)) {
        cfg->srcid = -1;
    }
}

void mcx_validatecfg(Config* cfg, float* detps, int dimdetps[2], int seedbyte) {
    if (!cfg->issrcfrom0) {
        cfg->srcpos.x--;
        cfg->srcpos.y--;
        cfg->srcpos.z--;

        for (unsigned int i = 0; i < cfg->extrasrclen; i++) {
            cfg->srcdata[i].srcpos.x--;
            cfg->srcdata[i].srcpos.y--;
            cfg->srcdata[i].srcpos.z--;
        }
    }

    if (cfg->tstart > cfg->tend || cfg->tstep == 0.f) {
        MCX_ERROR(-6, "incorrect time gate settings");
    }

    if (fabs(cfg->srcdir.x * cfg->srcdir.x + cfg->srcdir.y * cfg->srcdir.y + cfg->srcdir.z * cfg->srcdir.z - 1.f) > 1e-5) {
        MCX_ERROR(-6, "field 'srcdir' must be a unitary vector");
    }

    if (cfg->steps.x == 0.f || cfg->steps.y == 0.f || cfg->steps.z == 0.f) {
        MCX_ERROR(-6, "field 'steps' can not have zero elements");
    }

    if (cfg->tend <= cfg->tstart) {
        MCX_ERROR(-6, "field 'tend' must be greater than field 'tstart'");
    }

    int gates = (int)((cfg->tend - cfg->tstart) / cfg->tstep + 0.5);

    if (cfg->maxgate > (unsigned int)gates || cfg->maxgate == 0) {
        cfg->maxgate = gates;
    }

    if ((cfg->outputtype == otJacobian || cfg->outputtype == otWP || cfg->outputtype == otDCS ||
            cfg->outputtype == otRF || cfg->outputtype == otRFmus || cfg->outputtype == otWLTOF ||
            cfg->outputtype == otWPTOF) && cfg->seed != SEED_FROM_FILE) {
        MCX_ERROR(-6, "Jacobian output is only valid in the reply mode. Please define cfg.seed");
    }

    /* adjoint outputs and srcid -2: every detector becomes an extra disk source
       at its position and direction, radius detpos.w, and the field is stacked
       per source (src/mcx_utils.c:1902-1925).  A binding does this itself after
       validation (src/pmcxcl.cpp:1179-1198), so only a stand-alone run does it
       here. */
    if (cfg->parentid == mpStandalone && (MCX_IS_ADJOINT_TYPE(cfg->outputtype) || cfg->srcid == -2) &&
            cfg->seed != SEED_FROM_FILE && cfg->detnum > 0 && cfg->detdir != NULL) {
        mcx_append_detector_sources(cfg);
    }

    for (unsigned int i = 0; i < cfg->detnum; i++) {
        if (!cfg->issrcfrom0) {
            cfg->detpos[i].x--;
            cfg->detpos[i].y--;
            cfg->detpos[i].z--;
        }
    }

    if (cfg->shapedata && strstr(cfg->shapedata, ":") != NULL) {
        if (cfg->mediabyte > 4) {
            MCX_ERROR(-6, "rasterization of shapes must be used with label-based mediatype");
        }

        mcx_parse_shapestring(cfg, cfg->shapedata);
    }

    mcx_preprocess(cfg);

    cfg->his.maxmedia = cfg->medianum - 1;
    cfg->his.detnum = cfg->detnum;
    cfg->his.srcnum = cfg->srcnum;
    cfg->his.colcount = mcx_detreclen(cfg);
    cfg->his.savedetflag = cfg->savedetflag;
    cfg->his.totalsource = ((cfg->srcid <= 0) ? cfg->extrasrclen + 1 : 1);
    mcx_replayinit(cfg, detps, dimdetps, seedbyte);
}

/* writers: mcx_output.c */

/* ---------------- presets (reference src/mcx_bench.h, src/mcx_utils.c:5195-5203) ---------------- */
extern const char* mcx_preset_json(const char* name);

void mcx_loadbenchmark(char* key, Config* cfg) {
    const char* json = mcx_preset_json(key);

    if (!json) {
        MCX_ERROR(-1, "benchmark name is not recognized");
    }

    mcx_loadjson_string(json, cfg);
}

/* ---------------- command line (subset of reference src/mcx_utils.c:4608-5193) ---------------- */
void mcx_usage(Config* cfg, char* exename) {
    (void)cfg;
    printf("usage: %s <param1> <param2> ...\n"
           "  -f config     read a .json or .inp input file\n"
           "  --bench name  run a built-in benchmark (cube60, cube60b, cube60planar, cubesph60b, onecube, twocube)\n"
           "  -n photons    number of photons\n"
           "  -E seed       RNG seed, or a .mch/.jdat file of a -q 1 run to replay its photons\n"
           "  -b 0|1        Fresnel reflection at refractive-index mismatches\n"
           "  -B bcflags    per-face boundary condition letters [_ramc] and detection flags [01]\n"
           "  -s session    output name\n"
           "  -d 0|1        save detected photons\n"
           "  -H N          maximum detected photons\n"
           "  -q 0|1        save detected photon seeds\n"
           "  -O type       output type letter (x,f,e)\n"
           "  -w flags      detected photon fields [DSPMXVW]\n"
           "  -t N -T N     thread count and block size (disables autopilot)\n"
           "  -S 0|1        save fluence\n"
           "  -F format     output format: mc2, nii, hdr, ubj, tx3, jnii (default), bnii\n"
           "  -Z codec      JData compression: zlib (default), gzip, base64\n"
           "  --root path   output folder\n"
           "  -P N          replay the photons of detector N (0: all, -1: one output per detector)\n"
           "  -u mm         voxel edge length in mm\n"
           "  -r N          repeat the simulation N times\n"
           "  -z 0|1        source position counted from 0\n"
           "  -X 0|1 -x 0|1 save diffuse reflectance / exit position and direction\n"
           "  -V 0|1 -m 0|1 specular reflection at entry / save momentum transfer\n"
           "  --srcid N     simulate source N only (0: all together, -1: each separately)\n"
           "  --mathmode 0|1  0: hardware transcendentals, 1: bit-reproducible math\n",
           exename);
}

void mcx_parsecmd(int argc, char* argv[], Config* cfg) {
    char filename[MAX_PATH_LENGTH] = {0};
    char bench[64] = {0};
    int hasthread = 0;

    for (int i = 1; i < argc; i++) {
        const char* a = argv[i];
        const char* next = (i + 1 < argc) ? argv[i + 1] : NULL;
#define NEED_ARG() do { if (!next) { mcx_usage(cfg, argv[0]); MCX_ERROR(-1, "missing argument"); } i++; } while (0)

        if (!strcmp(a, "-h") || !strcmp(a, "--help")) {
            mcx_usage(cfg, argv[0]);
            exit(0);
        } else if (!strcmp(a, "-f") || !strcmp(a, "--input")) {
            NEED_ARG();
            snprintf(filename, sizeof(filename), "%s", next);
        } else if (!strcmp(a, "--bench") || !strcmp(a, "-Q")) {
            NEED_ARG();
            snprintf(bench, sizeof(bench), "%s", next);
        } else if (!strcmp(a, "-n") || !strcmp(a, "--photon")) {
            NEED_ARG();
            cfg->nphoton = (size_t)atof(next);
            flagset['n'] = 1;
        } else if (!strcmp(a, "-E") || !strcmp(a, "--seed")) {
            NEED_ARG();

            if (strstr(next, ".mch") || strstr(next, ".jdat")) {
                /* replay the photons saved in a detected-photon file */
                snprintf(cfg->seedfile, MAX_PATH_LENGTH, "%s", next);
                cfg->seed = SEED_FROM_FILE;
            } else {
                cfg->seed = atoi(next);
            }

            flagset['E'] = 1;
        } else if (!strcmp(a, "-b") || !strcmp(a, "--reflect")) {
            NEED_ARG();
            cfg->isreflect = (char)atoi(next);
            flagset['b'] = 1;
        } else if (!strcmp(a, "-B") || !strcmp(a, "--bc")) {
            NEED_ARG();
            memset(cfg->bc, 0, sizeof(cfg->bc));
            strncpy(cfg->bc, next, 12);
            flagset['B'] = 1;
        } else if (!strcmp(a, "-s") || !strcmp(a, "--session")) {
            NEED_ARG();
            snprintf(cfg->session, MAX_SESSION_LENGTH, "%s", next);
        } else if (!strcmp(a, "-d") || !strcmp(a, "--savedet")) {
            NEED_ARG();
            cfg->issavedet = (char)atoi(next);
            flagset['d'] = 1;
        } else if (!strcmp(a, "-H") || !strcmp(a, "--maxdetphoton")) {
            NEED_ARG();
            cfg->maxdetphoton = (unsigned int)atof(next);
            flagset['H'] = 1;
        } else if (!strcmp(a, "-q") || !strcmp(a, "--saveseed")) {
            NEED_ARG();
            cfg->issaveseed = (char)atoi(next);
            flagset['q'] = 1;
        } else if (!strcmp(a, "-O") || !strcmp(a, "--outputtype")) {
            NEED_ARG();
            char c = next[0];

            if (mcx_lookupindex(&c, outputtype)) {
                MCX_ERROR(-2, "the specified output data type is not recognized");
            }

            cfg->outputtype = c;
            flagset['O'] = 1;
        } else if (!strcmp(a, "-w") || !strcmp(a, "--savedetflag")) {
            NEED_ARG();
            cfg->savedetflag = isdigit((unsigned char)next[0]) ? (unsigned int)atoi(next) :
                               (unsigned int)mcx_parsedebugopt((char*)next, saveflag);
            flagset['w'] = 1;
        } else if (!strcmp(a, "-t") || !strcmp(a, "--thread")) {
            NEED_ARG();
            cfg->nthread = (unsigned int)atoi(next);
            hasthread = 1;
        } else if (!strcmp(a, "-T") || !strcmp(a, "--blocksize")) {
            NEED_ARG();
            cfg->nblocksize = (unsigned int)atoi(next);
            hasthread = 1;
        } else if (!strcmp(a, "-S") || !strcmp(a, "--save2pt")) {
            NEED_ARG();
            cfg->issave2pt = (char)atoi(next);
            flagset['S'] = 1;
        } else if (!strcmp(a, "-U") || !strcmp(a, "--normalize")) {
            NEED_ARG();
            cfg->isnormalized = (char)atoi(next);
            flagset['U'] = 1;
        } else if (!strcmp(a, "-G") || !strcmp(a, "--gpu")) {
            NEED_ARG();
            cfg->gpuid = atoi(next);
        } else if (!strcmp(a, "--mathmode")) {
            NEED_ARG();
            cfg->mathmode = atoi(next);
        } else if (!strcmp(a, "-e") || !strcmp(a, "--minenergy")) {
            NEED_ARG();
            cfg->minenergy = (float)atof(next);
            flagset['e'] = 1;
        } else if (!strcmp(a, "-F") || !strcmp(a, "--outputformat")) {
            NEED_ARG();
            int id = mcx_keylookup((char*)next, outputformat_names);

            if (id < 0) {
                MCX_ERROR(-2, "the specified output data format is not recognized");
            }

            cfg->outputformat = (char)id;
            flagset['F'] = 1;
        } else if (!strcmp(a, "-Z") || !strcmp(a, "--zip")) {
            NEED_ARG();
            cfg->zipid = isdigit((unsigned char)next[0]) ? atoi(next) : mcx_zipid(next);

            if (cfg->zipid < 0 || cfg->zipid > mzLz4hc) {
                MCX_ERROR(-2, "the specified compression method is not recognized");
            }
        } else if (!strcmp(a, "--root")) {
            NEED_ARG();
            snprintf(cfg->rootpath, MAX_PATH_LENGTH, "%s", next);
        } else if (!strcmp(a, "-P") || !strcmp(a, "--replaydet")) {
            NEED_ARG();
            cfg->replaydet = atoi(next);
        } else if (!strcmp(a, "-u") || !strcmp(a, "--unitinmm")) {
            NEED_ARG();
            cfg->unitinmm = (float)atof(next);
            flagset['u'] = 1;
        } else if (!strcmp(a, "-r") || !strcmp(a, "--repeat")) {
            NEED_ARG();
            cfg->respin = atoi(next);
        } else if (!strcmp(a, "-z") || !strcmp(a, "--srcfrom0")) {
            NEED_ARG();
            cfg->issrcfrom0 = (char)atoi(next);
        } else if (!strcmp(a, "-X") || !strcmp(a, "--saveref")) {
            NEED_ARG();
            cfg->issaveref = (char)atoi(next);
            flagset['X'] = 1;
        } else if (!strcmp(a, "-x") || !strcmp(a, "--saveexit")) {
            NEED_ARG();
            cfg->issaveexit = (char)atoi(next);
            flagset['x'] = 1;
        } else if (!strcmp(a, "-V") || !strcmp(a, "--specular")) {
            NEED_ARG();
            cfg->isspecular = (char)atoi(next);
            flagset['V'] = 1;
        } else if (!strcmp(a, "-m") || !strcmp(a, "--momentum")) {
            NEED_ARG();
            cfg->ismomentum = (char)atoi(next);
            flagset['m'] = 1;
        } else if (!strcmp(a, "--srcid")) {
            NEED_ARG();
            cfg->srcid = atoi(next);
        } else {
            mcx_usage(cfg, argv[0]);
            MCX_ERROR(-1, "unsupported command line option");
        }

#undef NEED_ARG
    }

    if (hasthread) {
        cfg->autopilot = 0;
    }

    if (bench[0]) {
        mcx_loadbenchmark(bench, cfg);
    } else if (filename[0]) {
        mcx_readconfig(filename, cfg);
    } else {
        mcx_usage(cfg, argv[0]);
        exit(0);
    }
}

/* ---------------- non-exiting wrappers ---------------- */
#define MCX_TRY(call)                      \
    jmp_buf jb;                            \
    void* old = mcx_push_errjmp(&jb);      \
    int rc = setjmp(jb);                   \
    if (rc == 0) {                         \
        mcx_errmsg[0] = '\0';              \
        call;                              \
    }                                      \
    mcx_pop_errjmp(old);                   \
    return rc;

int mcx_try_validatecfg(Config* cfg, float* detps, int dimdetps[2], int seedbyte) {
    MCX_TRY(mcx_validatecfg(cfg, detps, dimdetps, seedbyte));
}

int mcx_try_loadjson_string(const char* json, Config* cfg) {
    MCX_TRY(mcx_loadjson_string(json, cfg));
}

int mcx_try_loadbenchmark(const char* key, Config* cfg) {
    MCX_TRY(mcx_loadbenchmark((char*)key, cfg));
}

int mcx_try_readconfig(const char* fname, Config* cfg) {
    MCX_TRY(mcx_readconfig((char*)fname, cfg));
}

int mcx_try_parse_shapestring(Config* cfg, const char* shapedata) {
    MCX_TRY(mcx_parse_shapestring(cfg, shapedata));
}

int mcx_try_parsecmd(int argc, char* argv[], Config* cfg) {
    MCX_TRY(mcx_parsecmd(argc, argv, cfg));
}

int mcx_try_savedata(float* dat, size_t len, Config* cfg) {
    MCX_TRY(mcx_savedata(dat, len, cfg));
}

int mcx_try_savedetphoton(float* ppath, void* seeds, int count, int doappend, Config* cfg) {
    MCX_TRY(mcx_savedetphoton(ppath, seeds, count, doappend, cfg));
}

/* ---------------- ABI description ---------------- */
size_t mcx_config_sizeof(void) {
    return sizeof(Config);
}

#define F(name) {#name, offsetof(Config, name)}
static const struct {
    const char* name;
    size_t off;
} cfg_fields[] = {
    F(nphoton), F(nblocksize), F(nthread), F(seed), F(srcpos), F(srcdir), F(srciquv), F(tstart), F(tstep), F(tend),
    F(steps), F(dim), F(crop0), F(crop1), F(medianum), F(polmedianum), F(detnum), F(maxdetphoton), F(detradius),
    F(sradius), F(prop), F(polprop), F(detpos), F(detdir), F(smatrix), F(maxgate), F(respin), F(printnum), F(gpuid),
    F(vol), F(session), F(isrowmajor), F(isreflect), F(isref3), F(isrefint), F(isnormalized), F(issavedet),
    F(issave2pt), F(isgpuinfo), F(isspecular), F(isverbose), F(issrcfrom0), F(ismomentum), F(internalsrc),
    F(isdumpmask), F(issaveseed), F(istrajstokes), F(issaveexit), F(isatomic), F(issaveref), F(isdumpjson), F(zipid),
    F(srctype), F(autopilot), F(outputtype), F(outputformat), F(minenergy), F(unitinmm), F(omega), F(lambda),
    F(flog), F(his), F(energytot), F(energyabs), F(energyesc), F(rootpath), F(kernelfile), F(compileropt),
    F(shapedata), F(extrajson), F(clsource), F(maxvoidstep), F(voidtime), F(srcparam1), F(srcparam2), F(srcnum),
    F(srcpattern), F(deviceid), F(workload), F(exportfield), F(exportjacob), F(exportdetected), F(debuglevel),
    F(faststep), F(normalizer), F(replay), F(seeddata), F(replaydet), F(seedfile), F(jsonfile), F(maxjumpdebug),
    F(debugdatalen), F(gscatter), F(exportdebugdata), F(savedetflag), F(detectedcount), F(runtime), F(parentid),
    F(optlevel), F(mediabyte), F(bc), F(nphase), F(invcdf), F(nangle), F(angleinvcdf), F(srcid), F(extrasrclen),
    F(srcdata), F(mathmode), F(kernelms), F(jobms), F(runtimems), F(exportfield64), F(muamus), F(exportdetweight), F(replayinweight), F(exportfield32)
};
#undef F

long mcx_config_offsetof(const char* field) {
    for (size_t i = 0; i < sizeof(cfg_fields) / sizeof(cfg_fields[0]); i++) {
        if (!strcmp(cfg_fields[i].name, field)) {
            return (long)cfg_fields[i].off;
        }
    }

    return -1;
}

const char* mcx_version_string(void) {
    return MCX_VERSION;
}
